import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

DECKS = os.path.join(ROOT, "cudaprotoplanetarydisks_amd", "decks")
REFERENCE_IN = "/root/reference/in"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfargo_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running case")


BASE_DECK = {
    "AspectRatio": "0.05", "Sigma0": "6.3662e-4", "Viscosity": "1e-5", "SigmaSlope": "0.0",
    "FlaringIndex": "0.0", "PlanetConfig": os.path.join(DECKS, "jupiter_fixed.cfg"),
    "ThicknessSmoothing": "0.6", "Transport": "FARGO", "InnerBoundary": "EVANESCENT", "Disk": "YES",
    "OmegaFrame": "1.0", "Frame": "GUIDING-CENTER", "IndirectTerm": "YES", "Nrad": "128", "Nsec": "384",
    "Rmin": "0.4", "Rmax": "2.5", "RadialSpacing": "LOGARITHM", "Ntot": "199", "Ninterm": "20",
    "DT": "0.314159265359", "OutputDir": "out/", "SelfGravity": "NO", "Adiabatic": "NO", "Cooling": "NO",
}


def write_deck(path, **overrides):
    """Write a .par deck: the plain-FARGO workload with overrides
    (value None drops the key)."""
    d = dict(BASE_DECK)
    for k, v in overrides.items():
        if v is None:
            d.pop(k, None)
        else:
            d[k] = str(v)
    with open(path, "w") as f:
        f.write("# deck written by the test suite\n")
        for k, v in d.items():
            f.write(f"{k}\t\t{v}\n")
    return str(path)


def write_planets(path, rows):
    """rows: (name, dist, mass, accretion, feeldisk, feelothers)"""
    with open(path, "w") as f:
        f.write("# name dist mass accretion feeldisk feelothers\n")
        for r in rows:
            f.write("\t".join(str(x) for x in r) + "\n")
    return str(path)


@pytest.fixture
def deck(tmp_path):
    def make(name="d.par", **kw):
        kw.setdefault("OutputDir", str(tmp_path) + "/")
        return write_deck(tmp_path / name, **kw)
    return make
