"""Decks of the workloads this package is measured on.

BASE_DECK is the plain-FARGO disk: locally isothermal h = 0.05, flat
Sigma0 = 6.3662e-4, constant nu = 1e-5, one q = 1e-3 planet on a fixed
circular orbit at r = 1 (smoothing 0.6 H), FARGO transport, evanescent
edges, guiding-centre frame, 128 x 384 log grid over r in [0.4, 2.5].
Keys and value syntax are those of the reference's .par reader
(Interpret.cu:36-212).
"""
import os

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")

BASE_DECK = {
    "AspectRatio": "0.05", "Sigma0": "6.3662e-4", "Viscosity": "1e-5", "SigmaSlope": "0.0",
    "FlaringIndex": "0.0", "PlanetConfig": os.path.join(DECKS, "jupiter_fixed.cfg"),
    "ThicknessSmoothing": "0.6", "Transport": "FARGO", "InnerBoundary": "EVANESCENT", "Disk": "YES",
    "OmegaFrame": "1.0", "Frame": "GUIDING-CENTER", "IndirectTerm": "YES", "Nrad": "128", "Nsec": "384",
    "Rmin": "0.4", "Rmax": "2.5", "RadialSpacing": "LOGARITHM", "Ntot": "199", "Ninterm": "20",
    "DT": "0.314159265359", "OutputDir": "out/", "SelfGravity": "NO", "Adiabatic": "NO", "Cooling": "NO",
}

# seed and amplitude of the white-noise perturbation of the initial Sigma
NOISE_SEED = 20261202
NOISE_AMP = 1e-4


def write_deck(path, **overrides):
    """Write a .par deck: BASE_DECK with overrides (value None drops the key)."""
    d = dict(BASE_DECK)
    for k, v in overrides.items():
        if v is None:
            d.pop(k, None)
        else:
            d[k] = str(v)
    with open(path, "w") as f:
        f.write("# deck written by cudaprotoplanetarydisks_amd.workloads\n")
        for k, v in d.items():
            f.write(f"{k}\t\t{v}\n")
    return str(path)


def write_planets(path, rows):
    """rows: (name, dist, mass, accretion, feeldisk, feelothers), Psys.cu:93-142"""
    with open(path, "w") as f:
        f.write("# name dist mass accretion feeldisk feelothers\n")
        for r in rows:
            f.write("\t".join(str(x) for x in r) + "\n")
    return str(path)
