"""Parameter-sweep ensembles of isothermal disks on one GPU (fg_batch_*).

Draws (one seed for the whole ensemble): NRAD uniform over
{128, 192, 256, 384} with NSEC = 3 NRAD, log grid over r in [0.4, 2.5];
planet mass ratio q log-uniform in [1e-4, 3e-3]; alpha viscosity
log-uniform in [1e-4, 1e-2]; aspect ratio h uniform in [0.03, 0.08];
surface-density slope s uniform in [-1.5, 0], i.e. Sigma ~ r^s and the
deck's SigmaSlope = -s (Theo.cu:20 uses Sigma0 * r^-SIGMASLOPE); flaring
index 0 unless a range is given.  Every disk is the plain-FARGO setup
otherwise (one planet at r = 1 on a fixed circular orbit, evanescent
edges, guiding-centre frame) and runs in the FP32 bulk mode.
"""
import ctypes
import json
import os
import tempfile

import numpy as np

from . import native
from .native import call, lib
from .workloads import write_deck, write_planets

GRIDS = (128, 192, 256, 384)
TILES_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles.json")
_pd = ctypes.POINTER(ctypes.c_double)


def draw(n, seed, grids=GRIDS, flaring=(0.0, 0.0), nrad=None):
    """n disk parameter sets from one seed (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nr = int(nrad) if nrad else int(rng.choice(grids))
        d = {
            "nrad": nr, "nsec": 3 * nr,
            "q": float(np.exp(rng.uniform(np.log(1e-4), np.log(3e-3)))),
            "alpha": float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2)))),
            "h": float(rng.uniform(0.03, 0.08)),
            "slope": float(rng.uniform(-1.5, 0.0)),
            "flaring": float(rng.uniform(*flaring)) if flaring[1] > flaring[0] else float(flaring[0]),
        }
        out.append(d)
    return out


def write_disk_deck(directory, k, p):
    """Deck + planet file of disk k (Interpret.cu keys, Psys.cu columns)."""
    cfg = write_planets(os.path.join(directory, f"planet{k}.cfg"),
                        [("Planet", 1.0, repr(p["q"]), 0.0, "NO", "NO")])
    return write_deck(os.path.join(directory, f"disk{k}.par"), Nrad=p["nrad"], Nsec=p["nsec"],
                      AspectRatio=repr(p["h"]), SigmaSlope=repr(-p["slope"]),
                      FlaringIndex=repr(p["flaring"]), Viscosity=None, AlphaViscosity=repr(p["alpha"]),
                      PlanetConfig=cfg, OutputDir=os.path.join(directory, f"out{k}") + "/")


def load_tiles(path=TILES_JSON):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


class Ensemble:
    """fg_batch over n decks (FP32 bulk mode, one GPU)."""

    def __init__(self, decks, device=0, tiles=None, noise_seed=0, noise_amp=0.0):
        opt = native.FgSimOptions()
        opt.device = device
        opt.precision = native.FG_PREC_FP32
        opt.fused = 1
        opt.noise_seed = noise_seed
        opt.noise_amp = noise_amp
        self._decks = [os.fsencode(d) for d in decks]
        arr = (ctypes.c_char_p * len(decks))(*self._decks)
        h = ctypes.c_void_p()
        call("fg_batch_create", len(decks), arr, ctypes.byref(opt), ctypes.byref(h))
        self.h = h
        self.n = len(decks)
        self.shapes = []
        for k in range(self.n):
            s = self.sim_handle(k)
            v = ctypes.c_double()
            call("fg_sim_get", s, b"NRAD", ctypes.byref(v))
            nr = int(v.value)
            call("fg_sim_get", s, b"NSEC", ctypes.byref(v))
            self.shapes.append((nr + 1, int(v.value)))
        tiles = load_tiles() if tiles is None else tiles
        for k, (R, ns) in enumerate(self.shapes):
            t = tiles.get(f"{R - 1}x{ns}")
            if t:
                self.set_tiles(k, t["src_rows"], t["trad_band"])

    def close(self):
        if getattr(self, "h", None):
            lib().fg_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sim_handle(self, k):
        return ctypes.c_void_p(lib().fg_batch_sim(self.h, k))

    def set_tiles(self, k, src_rows, trad_band):
        call("fg_batch_set_tiles", self.h, k, int(src_rows), int(trad_band))

    def substep(self, t_end):
        na = ctypes.c_int()
        call("fg_batch_substep", self.h, float(t_end), ctypes.byref(na))
        return na.value

    def run(self, t_end):
        n = ctypes.c_long()
        call("fg_batch_run", self.h, float(t_end), ctypes.byref(n))
        return n.value

    def get(self, k, name):
        v = ctypes.c_double()
        call("fg_sim_get", self.sim_handle(k), name.encode(), ctypes.byref(v))
        return v.value

    def state(self, k):
        out = [np.empty(self.shapes[k]) for _ in range(3)]
        call("fg_sim_get_state", self.sim_handle(k), *[a.ctypes.data_as(_pd) for a in out], None)
        return out

    def dens(self, k):
        out = np.empty(self.shapes[k])
        call("fg_sim_get_state", self.sim_handle(k), out.ctypes.data_as(_pd), None, None, None)
        return out

    def ring(self, k, name):
        out = np.empty(self.shapes[k][0])
        call("fg_sim_ring", self.sim_handle(k), name.encode(), out.ctypes.data_as(_pd))
        return out

    def zones(self):
        return sum((R - 1) * ns for R, ns in self.shapes)


def make_ensemble(n, seed, directory=None, **kw):
    """Write the decks of n drawn disks and open them as one Ensemble."""
    directory = directory or tempfile.mkdtemp(prefix="ensemble_")
    params = draw(n, seed, **{k: v for k, v in kw.items() if k in ("grids", "flaring", "nrad")})
    decks = [write_disk_deck(directory, k, p) for k, p in enumerate(params)]
    ens_kw = {k: v for k, v in kw.items() if k in ("device", "tiles", "noise_seed", "noise_amp")}
    return Ensemble(decks, **ens_kw), params, directory


def main(argv=None):
    """Run a drawn ensemble to a target number of orbits and report the
    wall time, zone-updates and each disk's final ring-averaged Sigma."""
    import argparse
    import time
    ap = argparse.ArgumentParser(description="batched ensemble run")
    ap.add_argument("--disks", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--orbits", type=float, default=200.0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    ens, params, directory = make_ensemble(args.disks, args.seed)
    DT = 0.314159265359
    t_end = args.orbits * 2 * np.pi
    zones = [(R - 1) * ns for R, ns in ens.shapes]
    t0 = time.perf_counter()
    steps = 0
    last = t0
    while ens.substep(t_end):
        steps += 1
        now = time.perf_counter()
        if now - last > 30:
            done = sum(1 for k in range(ens.n) if ens.get(k, "PhysicalTime") >= t_end - 1e-9 * DT)
            print(json.dumps({"batch_substeps": steps, "seconds": round(now - t0, 1), "disks_done": done}),
                  flush=True)
            last = now
    wall = time.perf_counter() - t0
    zu = sum(zones[k] * ens.get(k, "substeps") for k in range(ens.n))
    report = {"disks": ens.n, "seed": args.seed, "orbits": args.orbits, "batch_substeps": steps,
              "seconds": round(wall, 2), "zones": int(sum(zones)), "zone_updates": int(zu),
              "zone_updates_per_s": round(zu / wall, 1)}
    finals = []
    for k in range(ens.n):
        dens, _, _ = ens.state(k)
        finals.append({"params": params[k], "time": ens.get(k, "PhysicalTime"),
                       "sigma_mean": dens[:-1].mean(axis=1).tolist()})
    print(json.dumps(report), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"report": report, "disks": finals}, f)
    ens.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
