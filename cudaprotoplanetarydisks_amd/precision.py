"""FP32 bulk mode against the FP64 path: per-field error report.

Runs the same deck twice on one GPU, FP64 (the path that reproduces the
reference's arithmetic) and FP32 bulk, from the same seeded initial state,
advancing both one DT (one AlgoGas call, SourceEuler.cu:199-307) at a time
and comparing the fields every `every` DTs:

    linf = max|a32 - a64| / max|a64|,   l2 = ||a32 - a64||_2 / ||a64||_2

over the active rings (0..NRAD-1), for Sigma, vrad and vtheta, with the cell
of each largest difference.  With --b-precision fp64 --b-amp-rel E the
second run is FP64 too, its initial noise amplitude scaled by 1 + E: how far
the flow itself carries a small initial difference.

    python -m cudaprotoplanetarydisks_amd.precision --nrad 1024 --nsec 3072 --orbits 50
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

from .native import call
from .sim import Simulation
from .workloads import NOISE_AMP, NOISE_SEED, write_deck

FIELDS = ("dens", "vrad", "vtheta")


def field_errors(a, b, nr):
    """Errors of a against b; for each field also the cell (ring, sector) of
    the largest difference, the field's value there relative to its maximum,
    and the cell of that maximum."""
    out = {}
    for name, x, y in zip(FIELDS, a, b):
        x = x[:nr]
        y = y[:nr]
        d = x - y
        ymax = max(np.abs(y).max(), 1e-300)
        at = np.unravel_index(int(np.abs(d).argmax()), d.shape)
        out[name] = {"linf": float(np.abs(d).max() / ymax),
                     "l2": float(np.sqrt((d * d).sum() / max((y * y).sum(), 1e-300))),
                     "at": [int(at[0]), int(at[1])], "value_at": float(y[at] / ymax),
                     "max_at": [int(v) for v in np.unravel_index(int(np.abs(y).argmax()), y.shape)]}
    return out


def compare(deck, n_dt, every=1, log=None, seed=NOISE_SEED, amp=NOISE_AMP, b_precision="fp32", b_amp_rel=0.0):
    """Advance an FP64 run and a second run (FP32 bulk, or FP64 with the
    initial noise amplitude changed by the factor 1 + b_amp_rel, which
    measures how far the flow itself amplifies a small initial difference)
    of `deck` n_dt DTs; yield a report every `every` DTs (and after the
    last)."""
    s64 = Simulation(deck, fused=True, precision="fp64", noise_seed=seed, noise_amp=amp)
    s32 = Simulation(deck, fused=True, precision=b_precision, noise_seed=seed, noise_amp=amp * (1.0 + b_amp_rel))
    nr = s64.nrad
    DT = s64.get("DT")
    t64 = t32 = 0.0
    try:
        for k in range(1, n_dt + 1):
            t0 = time.perf_counter()
            s64.algogas()
            call("fg_sync", s64.ctx.h)
            t1 = time.perf_counter()
            s32.algogas()
            call("fg_sync", s32.ctx.h)
            t2 = time.perf_counter()
            t64 += t1 - t0
            t32 += t2 - t1
            if k % every == 0 or k == n_dt:
                rep = {"dt_index": k, "orbits": k * DT / (2 * math.pi),
                       "errors": field_errors(s32.state(), s64.state(), nr),
                       "seconds_fp64": round(t64, 3), "seconds_fp32": round(t32, 3)}
                if log:
                    log(rep)
                yield rep
    finally:
        s64.close()
        s32.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--nrad", type=int, default=1024)
    ap.add_argument("--nsec", type=int, default=3072)
    ap.add_argument("--orbits", type=float, default=50.0)
    ap.add_argument("--out", default=None, help="write the JSON lines here too")
    ap.add_argument("--b-precision", default="fp32", choices=("fp32", "fp64"),
                    help="precision of the run compared with the FP64 one")
    ap.add_argument("--b-amp-rel", type=float, default=0.0,
                    help="relative change of that run's initial noise amplitude")
    args = ap.parse_args(argv)
    with tempfile.TemporaryDirectory() as tmp:
        deck = write_deck(os.path.join(tmp, "precision.par"), Nrad=args.nrad, Nsec=args.nsec,
                          OutputDir=tmp + "/")
        DT = float(0.314159265359)
        per_orbit = max(1, round(2 * math.pi / DT))
        n_dt = int(round(args.orbits * 2 * math.pi / DT))
        fout = open(args.out, "w") if args.out else None

        def log(rep):
            line = json.dumps(rep)
            print(line, flush=True)
            if fout:
                fout.write(line + "\n")
                fout.flush()

        worst = {f: 0.0 for f in FIELDS}
        for rep in compare(deck, n_dt, every=per_orbit, log=log, b_precision=args.b_precision,
                           b_amp_rel=args.b_amp_rel):
            for f in FIELDS:
                worst[f] = max(worst[f], rep["errors"][f]["linf"])
        summary = {"grid": [args.nrad, args.nsec], "orbits": args.orbits, "b_precision": args.b_precision,
                   "b_amp_rel": args.b_amp_rel, "worst_linf": worst}
        log(summary)
        if fout:
            fout.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
