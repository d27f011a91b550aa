"""Pins of the oracle's diagnostics: ComputeForce with the Hill cut
(Force.cu:102-135, Kernels.cu:693-729), CircumPlanetaryMass
(SourceEuler.cu:549-563, Kernels.cu:236-251) and the layouts of the files
UpdateLog (Force.cu:19-84) and WriteBigPlanetSystemFile (Output.cu:62-107)
append.  The reference ships these calls commented out (Main.cu:302,
314-315; SourceEuler.cu:297-298), so nothing it holds pins them; the
checks below come from what the mathematics fixes."""
import math
import os

import numpy as np
import pytest

from oracle_binding import Oracle

DIMFXY = 11


@pytest.fixture
def orc(deck):
    o = Oracle(deck(Nrad=24, Nsec=60))
    rng = np.random.default_rng(3)
    d = o.view("Dens")
    d[:-1] = rng.uniform(1e-4, 1e-3, d[:-1].shape)
    yield o
    o.close()


def parts(gf, n=DIMFXY):
    """GlobalForce -> (fx_inner, fx_outer, fy_inner, fy_outer), each [k]."""
    return gf[:n], gf[n:2 * n], gf[2 * n:3 * n], gf[3 * n:]


def test_tiny_hill_radius_leaves_every_variant_uncut(orc):
    # d^2 / rh_k^2 >= 0.16 / 1e-24 for every cell: exp underflows to 0,
    # hill_cut is exactly 1 and every k reproduces k = 0 bit for bit
    gf = orc.force_hill(0.0, 0.0, 0.01, DIMFXY, 1.2, 1e-12)
    for p in parts(gf):
        assert np.all(p == p[0])


def test_split_radius_zero_puts_every_cell_outside(orc):
    # a = 0: Rmed < a never holds; the outer sums are ComputeAccel's
    gf = orc.force_hill(0.7, -0.4, 0.02, DIMFXY, 0.0, 1e-12)
    fxi, fxo, fyi, fyo = parts(gf)
    assert np.all(fxi == 0.0) and np.all(fyi == 0.0)
    acc = orc.compute_accel(0.7, -0.4, 0.02)
    assert fxo[0] == acc[0] and fyo[0] == acc[1]


def test_split_radius_beyond_the_disk_puts_every_cell_inside(orc):
    gf = orc.force_hill(0.7, -0.4, 0.02, DIMFXY, 10.0, 1e-12)
    fxi, fxo, fyi, fyo = parts(gf)
    assert np.all(fxo == 0.0) and np.all(fyo == 0.0)
    acc = orc.compute_accel(0.7, -0.4, 0.02)
    assert fxi[0] == acc[0] and fyi[0] == acc[1]


def test_inner_sum_is_the_force_of_the_rings_below_a(orc):
    # Rmed[i] < a is strict: with a = Rmed[9], rings 0..8 are inner.  The
    # inner sums equal ComputeAccel on a density with rings >= 9 removed
    # (a zero cell adds +-0 in either form), the outer ones the complement.
    rmed = orc.field("Rmed")
    dens = orc.field("Dens")
    x, y, rs = 1.1, 0.3, 0.04
    gf = orc.force_hill(x, y, rs, DIMFXY, rmed[9], 1e-12)
    fxi, fxo, fyi, fyo = parts(gf)
    d = orc.view("Dens")
    d[9:-1] = 0.0
    acc_in = orc.compute_accel(x, y, rs)
    d[:] = dens
    d[:9] = 0.0
    acc_out = orc.compute_accel(x, y, rs)
    d[:] = dens
    assert (fxi[0], fyi[0]) == tuple(acc_in)
    assert (fxo[0], fyo[0]) == tuple(acc_out)


def test_single_cell_matches_the_tapered_point_mass_force(orc):
    # one massive cell: each sum is that cell's pull on (x, y), the
    # smoothed inverse-square force times 1 - exp(-d^2 / (rh k/10)^2)
    d = orc.view("Dens")
    d[:] = 0.0
    i0, j0 = 13, 17
    d[i0, j0] = 2e-3
    rmed, surf = orc.field("Rmed"), orc.field("Surf")
    th = 2 * math.pi * j0 / orc.nsec
    cx, cy = rmed[i0] * math.cos(th), rmed[i0] * math.sin(th)
    x, y, rs, a, rh = cx + 0.05, cy - 0.08, 0.03, 0.5, 0.2
    gf = orc.force_hill(x, y, rs, DIMFXY, a, rh)
    fxi, fxo, fyi, fyo = parts(gf)
    assert np.all(fxi == 0.0)  # the cell is outside r = a
    m = surf[i0] * 2e-3
    dx, dy = cx - x, cy - y
    d2 = dx * dx + dy * dy
    f = m / (d2 + rs * rs) ** 1.5
    for k in range(DIMFXY):
        cut = 1.0 if k == 0 else 1.0 - math.exp(-d2 / (rh * k / (DIMFXY - 1)) ** 2)
        assert fxo[k] == pytest.approx(f * dx * cut, rel=1e-13, abs=0)
        assert fyo[k] == pytest.approx(f * dy * cut, rel=1e-13, abs=0)


def test_circumplanetary_mass_closed_forms(orc):
    d = orc.view("Dens")
    d[:] = 1.0
    rmin, rmax = 0.4, 2.5
    # a radius that holds the whole disk: the disk's area, pi (Rmax^2 - Rmin^2)
    assert orc.circumplanetary_mass(0.0, 0.0, 10.0) == pytest.approx(math.pi * (rmax ** 2 - rmin ** 2), rel=1e-13)
    assert orc.circumplanetary_mass(1.0, 0.0, 0.0) == 0.0
    # centred on the star, radius between Rmed[6] and Rmed[7]: rings 0..6
    rmed, surf = orc.field("Rmed"), orc.field("Surf")
    hill = 0.5 * (rmed[6] + rmed[7])
    want = math.fsum(surf[:7] * orc.nsec)
    assert orc.circumplanetary_mass(0.0, 0.0, hill) == pytest.approx(want, rel=1e-14)


def test_circumplanetary_mass_counts_cells_strictly_inside(orc):
    # a radius equal to a cell's distance leaves that cell out (dist < R)
    d = orc.view("Dens")
    d[:] = 0.0
    d[5, 0] = 1.0
    rmed, surf = orc.field("Rmed"), orc.field("Surf")
    assert orc.circumplanetary_mass(0.0, 0.0, rmed[5]) == 0.0
    assert orc.circumplanetary_mass(0.0, 0.0, np.nextafter(rmed[5], 1.0)) == surf[5]


def read_table(path):
    return np.loadtxt(path, ndmin=2)


def test_log_files_layout(deck, tmp_path):
    out = str(tmp_path / "o") + "/"
    os.makedirs(out)
    path = deck(Nrad=24, Nsec=60, Ninterm=2, Ntot=5, OutputDir=out)
    o = Oracle(path, write_files=True)
    o.enable_diagnostics()
    mdcp = o.get("mdcp")
    assert mdcp > 0.0
    o.run(0, 5)
    tq = read_table(out + "tqwk0.dat")
    big = read_table(out + "bigplanet0.dat")
    pl = read_table(out + "planet0.dat")
    # one line per main-loop iteration, stamped with the output number of
    # the iteration before (Main.cu:312-318)
    assert list(tq[:, 0]) == [0, 0, 0, 1, 1, 2]
    assert list(big[:, 0]) == [0, 0, 0, 1, 1, 2]
    assert tq.shape[1] == 10 and big.shape[1] == 11
    # the time column advances by DT per iteration
    assert np.allclose(np.diff(tq[:, 9]), 0.314159265359, rtol=1e-12)
    # k = 0 is the uncut force, k = dimfxy - 1 the one the ex_ columns use
    k0 = read_table(out + "tqwk0_0.dat")
    k10 = read_table(out + "tqwk0_10.dat")
    assert np.array_equal(k0[:, 1:3], tq[:, 1:3]) and np.array_equal(k0[:, 3:5], tq[:, 5:7])
    assert np.array_equal(k10[:, 1:3], tq[:, 3:5]) and np.array_equal(k10[:, 3:5], tq[:, 7:9])
    # planet lines carry the initial mass and the excess of the last substep
    assert np.all(pl[:, 9] == mdcp)
    assert pl[0, 10] == 0.0
    assert np.all(big[1:, 10] == big[1:, 9] - mdcp)
    o.close()
