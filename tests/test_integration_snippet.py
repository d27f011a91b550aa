"""The reference-side binding shown in INTEGRATION.md section 2.

CPU: the snippet in INTEGRATION.md is exactly the AlgoGas replacement in
tests/integration_snippet.c, it compiles against include/fargo_hip.h and
links libfargo_hip.so, and every library call in it names a declared entry
point with the declared number of arguments.
GPU: run through the C-ABI, it matches the library's own host driver
(fg_sim_substep) bit for bit on an isothermal and an adiabatic deck."""
import os
import re
import subprocess

import pytest

from conftest import ROOT, write_planets

SRC = os.path.join(ROOT, "tests", "integration_snippet.c")
LIBDIR = os.path.join(ROOT, "cudaprotoplanetarydisks_amd", "lib")
BIN = os.path.join(LIBDIR, "integration_snippet")
HEADER = os.path.join(ROOT, "include", "fargo_hip.h")


def snippet_from(text, begin, end):
    i = text.index(begin) + len(begin)
    return text[i:text.index(end, i)].strip("\n")


def c_snippet():
    return snippet_from(open(SRC).read(), "/* ---- snippet ", "/* ---- end snippet")


def doc_snippet():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = snippet_from(doc, "<!-- snippet: tests/integration_snippet.c -->\n```c\n", "```")
    return block


def build_binary():
    """gcc: the maintainer's side is plain C against the header."""
    if os.path.exists(BIN) and os.path.getmtime(BIN) >= os.path.getmtime(SRC):
        return BIN
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Werror", "-ffp-contract=off", "-fno-builtin-sin",
                    "-fno-builtin-cos", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", BIN, f"-L{LIBDIR}",
                    "-lfargo_hip", "-lm", f"-Wl,-rpath,{LIBDIR}"], check=True, capture_output=True, text=True)
    return BIN


def test_integration_md_shows_the_tested_snippet():
    c = c_snippet()
    body = c[c.index("\n") + 1:]  # drop the marker line's tail
    assert body.strip() in doc_snippet(), "INTEGRATION.md section 2 differs from tests/integration_snippet.c"


def test_snippet_calls_match_the_header():
    header = open(HEADER).read()
    decl = {m.group(1): m.group(2) for m in re.finditer(r"fg_status (fg_\w+)\(([^;]*?)\);", header, re.S)}
    calls = re.findall(r"\b(fg_\w+)\(([^;{]*?)\)\)?;", doc_snippet())
    assert len(calls) >= 12
    seen = set()
    for name, args in calls:
        assert name in decl, f"{name} is not declared in fargo_hip.h"
        nargs = len([a for a in re.split(r",(?![^(]*\))", args) if a.strip()])
        assert nargs == len(decl[name].split(",")), f"{name}: {nargs} arguments, header declares {decl[name]}"
        seen.add(name)
    # every routine of SourceEuler.cu:228-295 is bound
    assert {"fg_cfl", "fg_disk_force", "fg_potential", "fg_correct_vtheta", "fg_boundary", "fg_pressure",
            "fg_substep1", "fg_substep2", "fg_substep3", "fg_transport", "fg_temperature", "fg_sound_speed"} <= seen


def test_snippet_compiles_and_links():
    if not os.path.exists(os.path.join(LIBDIR, "libfargo_hip.so")):
        pytest.skip("libfargo_hip.so not built")
    assert os.access(build_binary(), os.X_OK)


@pytest.mark.gpu
@pytest.mark.parametrize("adiabatic", [False, True], ids=["isothermal", "adiabatic-cooling"])
def test_snippet_matches_fg_sim_substep_bitwise(tmp_path, deck, adiabatic):
    cfg = write_planets(tmp_path / "p.cfg", [("P", 1.0, 1e-3, 0.0, "YES", "NO")])
    kw = dict(Nrad=32, Nsec=64, Frame="FIXED", OmegaFrame=0.0, PlanetConfig=cfg, MassTaper=3.0,
              ThicknessSmoothing=0.6, AspectRatio=0.05, FlaringIndex=0.0)
    if adiabatic:
        kw.update(Adiabatic="YES", Cooling="YES", AdiabaticIndex=1.4, CoolingTime0=62.8)
    path = deck(**kw)
    r = subprocess.run([build_binary(), path, "3", "3.0", "0.6", "0.05", "0.0"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout + r.stderr)
    assert r.returncode == 0
    assert "bit for bit over 3 AlgoGas calls" in r.stdout
