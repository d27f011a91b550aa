"""Tile-shape autotuner of the FP32 fused kernels, per grid size.

For each (NRAD, NSEC) it builds a batch of `copies` disks of that grid (an
ensemble's worth of parallel work: with one small disk the GPU is mostly
idle and small tiles win for the wrong reason) and times batch substeps
(the whole captured graph: every phase and reduction) for every candidate
(rings per source tile) x (rings per radial-sweep band), keeping the
fastest.  Results go to a JSON cache keyed "NRADxNSEC" that
Ensemble reads.  Tile shapes change only the work decomposition: results
are identical for every candidate.

    python -m cudaprotoplanetarydisks_amd.autotune --grids 128x384 256x768
"""
import argparse
import json
import os
import sys
import tempfile
import time

from .ensemble import TILES_JSON, Ensemble, load_tiles
from .workloads import write_deck

SRC_ROWS = (8, 16, 32)
TRAD_BANDS = (8, 16, 32, 64)


def time_candidate(ens, rows, band, warmup, steps):
    for k in range(ens.n):
        ens.set_tiles(k, rows, band)
    t_end = 1e30
    for _ in range(warmup):
        ens.substep(t_end)
    t0 = time.perf_counter()
    for _ in range(steps):
        ens.substep(t_end)
    return (time.perf_counter() - t0) / steps * 1e3


def tune_grid(nr, ns, warmup=5, steps=30, log=None, copies=16):
    with tempfile.TemporaryDirectory() as tmp:
        deck = write_deck(os.path.join(tmp, "tune.par"), Nrad=nr, Nsec=ns, Viscosity=None, AlphaViscosity=1e-3,
                          OutputDir=tmp + "/")
        with Ensemble([deck] * copies, tiles={}) as ens:
            results = []
            for rows in SRC_ROWS:
                for band in TRAD_BANDS:
                    ms = time_candidate(ens, rows, band, warmup, steps)
                    results.append((ms, rows, band))
                    if log:
                        log(f"{nr}x{ns} rows={rows} band={band}: {ms:.4f} ms/substep")
    ms, rows, band = min(results)
    return {"src_rows": rows, "trad_band": band, "ms_per_substep": round(ms, 5),
            "candidates": [{"src_rows": r, "trad_band": b, "ms": round(m, 5)} for m, r, b in results]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--grids", nargs="+", default=["128x384", "192x576", "256x768", "384x1152"])
    ap.add_argument("--out", default=TILES_JSON)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--copies", type=int, default=16, help="disks of the grid per timed batch")
    args = ap.parse_args(argv)
    cache = load_tiles(args.out)
    for g in args.grids:
        nr, ns = (int(x) for x in g.split("x"))
        cache[f"{nr}x{ns}"] = tune_grid(nr, ns, steps=args.steps, log=lambda m: print(m, flush=True),
                                        copies=args.copies)
        print(json.dumps({g: {k: v for k, v in cache[f'{nr}x{ns}'].items() if k != 'candidates'}}), flush=True)
    with open(args.out, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
