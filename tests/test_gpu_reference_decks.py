"""Every deck the reference ships (in/*.par with its in/*.cfg planet table),
at a reduced grid, against the oracle substep by substep.  Together they
cover the open, non-reflecting and evanescent inner edges, the fixed,
corotating and guiding-centre frames, arithmetic and logarithmic radii,
alpha viscosity with flaring, adiabatic runs with and without cooling,
forced circular orbits and several planets that feel the disk.

The settings are restated here as data (the reference tree is not present
where these tests run); only Nrad/Nsec are reduced so the serial oracle
finishes in seconds.  Adiabatic substeps use the device exp/log
(Kernels.cu:56-115), hence the 1e-12 relative bar of the adiabatic cases.
"""
import pytest

from conftest import write_planets
from test_gpu_parity import run_pair

pytestmark = pytest.mark.gpu

# in/*.cfg: (name, distance, mass ratio, accretion, feel disk, feel others)
PLANETS = {
    "Jup.cfg": [("Jupiter", 1.0, 0.001, 0.0, "NO", "NO")],
    "Jup2.cfg": [("Jupiter", 4.5, 0.001, 0.0, "YES", "NO"), ("Jupiter2", 6.5, 0.001, 0.0, "YES", "NO"),
                 ("Jupiter3", 3.5, 0.001, 0.0, "YES", "NO"), ("Jupiter4", 5.2, 0.001, 0.0, "YES", "NO")],
    "2p.cfg": [("P", 12.0, 4.16e-3, 0.0, "No", "No"), ("PP", 53.0, 2.08e-3, 0.0, "No", "No")],
    "adsg.cfg": [("MyPlanet", 1.0, 2.15786376476132e-05, 0.0, "NO", "NO")],
}

_ADIABATIC = dict(SelfGravity="NO", Adiabatic="YES", AdiabaticIndex=1.4)
_GC = dict(Transport="FARGO", Disk="YES", OmegaFrame=1.0, Frame="GUIDING-CENTER", IndirectTerm="YES",
           ThicknessSmoothing=0.6, AspectRatio=0.05, Viscosity=1e-5, SigmaSlope=0.0, FlaringIndex=0.0,
           DT=0.314159265359)

# in/<deck>.par: key settings (Nrad, Nsec reduced from the deck's)
DECKS = {
    "adsg": ("adsg.cfg", dict(AspectRatio=0.05, Sigma0=2e-3, Viscosity=1e-5, SigmaSlope=0.5, FlaringIndex=0.0,
                              Cooling="YES", CoolingTime0=62.8, ForcedCircular="YES", ThicknessSmoothing=0.6,
                              Transport="FARGO", InnerBoundary="EVANESCENT", Disk="YES", OmegaFrame=0.0,
                              Frame="COROTATING", IndirectTerm="YES", Nrad=40, Nsec=120, Rmin=0.4, Rmax=1.8,
                              RadialSpacing="A", DT=0.314159265359, **_ADIABATIC)),
    "hd100": ("2p.cfg", dict(AspectRatio=0.0297109, Sigma0=0.83e-6, SigmaSlope=1.0, FlaringIndex=0.142857,
                             Viscosity=None, AlphaViscosity=1e-4, Cooling="NO", ForcedCircular="NO",
                             ThicknessSmoothing=0.6, Transport="FARGO", Nrad=50, Nsec=150, Rmin=1.0, Rmax=80.0,
                             InnerBoundary="EVANESCENT", RadialSpacing="A", Disk="YES", OmegaFrame=0.0,
                             Frame="FIXED", IndirectTerm="YES", DT=15.89534122527376, **_ADIABATIC)),
    "temp": ("Jup.cfg", dict(_GC, Sigma0=6.3661977237e-4, InnerBoundary="NONREFLECTING", Nrad=48, Nsec=48,
                             Rmin=3.2, Rmax=7.2, RadialSpacing="LOGARITHM", Cooling="NO", CoolingTime0=62.8,
                             **_ADIABATIC)),
    "template": ("Jup2.cfg", dict(_GC, Sigma0=5.1e-5, InnerBoundary="OPEN", Nrad=32, Nsec=96, Rmin=3.2, Rmax=7.2,
                                  RadialSpacing="LOGARITHM", Cooling="YES", CoolingTime0=62.8, WriteQplus="YES",
                                  **_ADIABATIC)),
    "template2": ("Jup.cfg", dict(_GC, Sigma0=6.3661977237e-4, InnerBoundary="EVANESCENT", Nrad=32, Nsec=96,
                                  Rmin=0.4, Rmax=2.5, RadialSpacing="A", Cooling="YES", CoolingTime0=62.8,
                                  WriteQplus="YES", **_ADIABATIC)),
    # template3 and template4 differ only in grid size and output cadence
    "template3_4": ("Jup.cfg", dict(_GC, Sigma0=5.1e-5, InnerBoundary="NONREFLECTING", Nrad=48, Nsec=96,
                                    Rmin=0.4, Rmax=2.5, RadialSpacing="LOGARITHM", Cooling="NO",
                                    CoolingTime0=62.8, **_ADIABATIC)),
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_reference_deck_matches_oracle(deck, tmp_path, name):
    cfg_name, kw = DECKS[name]
    cfg = write_planets(tmp_path / cfg_name, PLANETS[cfg_name])
    path = deck(f"{name}.par", PlanetConfig=cfg, **kw)
    run_pair(path, 24, 1e-12, adiabatic=True)
