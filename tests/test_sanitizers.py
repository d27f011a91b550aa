"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only; GPU sanitizers are not available on this pool): a few
substeps of every code path the GPU tests compare against -- isothermal
and adiabatic decks, open / non-reflecting / evanescent edges, rotating
frames, alpha viscosity, self-gravity, the diagnostics and the outputs --
must run without a sanitizer report."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, write_planets

ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    out = tmp_path_factory.mktemp("asan") / "sanitize_driver"
    cmd = ["gcc", "-O1", "-g", "-std=gnu99", "-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           f"-I{ORACLE}", os.path.join(ORACLE, "fargo_oracle.c"), os.path.join(ROOT, "tests", "sanitize_driver.c"),
           "-o", str(out), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr and "cannot find" in r.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


CASES = {
    "isothermal": dict(Nrad=24, Nsec=48),
    "adiabatic_cooling": dict(Nrad=24, Nsec=48, Adiabatic="YES", AdiabaticIndex=1.4, Cooling="YES",
                              CoolingTime0=62.8),
    "open_inner": dict(Nrad=20, Nsec=40, InnerBoundary="OPEN"),
    "nonreflecting": dict(Nrad=20, Nsec=40, InnerBoundary="NONREFLECTING"),
    "corotating_alpha": dict(Nrad=20, Nsec=40, Frame="COROTATING", Viscosity=None, AlphaViscosity=1e-3),
    "selfgravity": dict(Nrad=16, Nsec=24, SelfGravity="YES", ThicknessSmoothing=0.3),
    "standard_transport": dict(Nrad=16, Nsec=32, Transport="STANDARD"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_runs_clean_under_asan_ubsan(driver, deck, tmp_path, case):
    kw = dict(CASES[case])
    cfg = write_planets(tmp_path / "p.cfg", [("A", 1.0, 1e-3, 0.0, "YES", "YES"), ("B", 1.5, 3e-4, 0.0, "YES", "YES")])
    path = deck(PlanetConfig=cfg, **kw)
    # verify_asan_link_order=0: the environment may preload a library of its
    # own ahead of the ASan runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=97:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([driver, path, str(out) + "/", "6", "1"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.startswith("ok 6 substeps")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
