"""Learned warm start for parameter sweeps.

An MLP (5 inputs, 4 hidden layers of 256 with GELU, 512 outputs; BF16
weights, FP32 accumulation, trained on the GPU by fg_mlp_*) maps a disk's
parameters (log q, log alpha, h, Sigma slope, flaring) to its
ring-averaged profiles after `orbits` orbits from the plain initial state:
Sigma(r) / Sigma_init(r) - 1 and v_r(r) / (alpha h^2) at 256 log-spaced
radii in [0.4, 2.5].  The profiles come from the batched solver itself
(fg_batch_*).

Warm start: the predicted Sigma(r) and v_r(r), interpolated to the disk's
rings, with v_theta(r) in radial equilibrium with the predicted pressure
gradient, replace the initial state through fg_sim_set_state (the restart
path); the plain method then continues from there.  Both starts are
measured by the same steadiness: the change of the orbit-averaged,
ring-averaged Sigma from one orbit to the next, mass-weighted over the
rings outside the damping zones (sum |dSigma| / sum Sigma per orbit), and
the time to steady is the first orbit where it falls below eps.

    python -m cudaprotoplanetarydisks_amd.emulator study --n-train 512 --n-eval 32
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

from .ensemble import Ensemble, write_disk_deck
from .mlp import MLP

NOUT = 256
R_OUT = np.geomspace(0.4, 2.5, NOUT)
WIDTHS = [5, 256, 256, 256, 256, 2 * NOUT]
ORBIT = 2.0 * math.pi
DT = 0.314159265359
RANGES = {"lq": (math.log(1e-4), math.log(3e-3)), "la": (math.log(1e-4), math.log(1e-2)),
          "h": (0.03, 0.08), "slope": (-1.5, 0.0), "flaring": (0.0, 0.25)}


def draw(n, seed, nrad):
    """Disk parameters as the ensemble draws them, plus flaring in [0, 0.25]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append({"nrad": nrad, "nsec": 3 * nrad,
                    "q": float(np.exp(rng.uniform(*RANGES["lq"]))),
                    "alpha": float(np.exp(rng.uniform(*RANGES["la"]))),
                    "h": float(rng.uniform(*RANGES["h"])),
                    "slope": float(rng.uniform(*RANGES["slope"])),
                    "flaring": float(rng.uniform(*RANGES["flaring"]))})
    return out


def features(params):
    """Inputs scaled to [-1, 1]."""
    def s(v, key):
        lo, hi = RANGES[key]
        return 2.0 * (v - lo) / (hi - lo) - 1.0
    return np.array([[s(math.log(p["q"]), "lq"), s(math.log(p["alpha"]), "la"), s(p["h"], "h"),
                      s(p["slope"], "slope"), s(p["flaring"], "flaring")] for p in params], np.float32)


def vr_scale(p):
    return p["alpha"] * p["h"] ** 2


def ring_profiles(ens, k):
    """Ring-averaged Sigma / Sigma_init and v_r of disk k (active rings)."""
    dens, vr, _ = ens.state(k)
    nr = dens.shape[0] - 1
    sig = dens[:nr].mean(axis=1) / ens.ring(k, "SigmaMed")[:nr]
    return sig, vr[:nr].mean(axis=1)


def to_output(ens, k, p):
    sig, vr = ring_profiles(ens, k)
    rmed = ens.ring(k, "Rmed")[:len(sig)]
    rinf = ens.ring(k, "Rinf")[:len(sig)]
    s = np.interp(np.log(R_OUT), np.log(rmed), sig) - 1.0
    v = np.interp(np.log(R_OUT), np.log(np.maximum(rinf, 1e-12)), vr) / vr_scale(p)
    return np.concatenate([s, v]).astype(np.float32)


def steadiness(prev, cur, rmed, rlo=0.6, rhi=2.0):
    """Relative change of the orbit-averaged, ring-averaged Sigma between
    consecutive orbits, mass-weighted over rlo < r < rhi (outside the
    damping zones): sum |cur - prev| / sum prev."""
    m = (rmed > rlo) & (rmed < rhi)
    return float(np.abs(cur[m] - prev[m]).sum() / np.abs(prev[m]).sum())


def open_ensemble(params, directory, tiles=None):
    decks = [write_disk_deck(directory, k, p) for k, p in enumerate(params)]
    return Ensemble(decks, tiles=tiles)


def generate(params, orbits, chunk=128, log=None):
    """Run every disk `orbits` orbits from the plain initial state (batched,
    `chunk` disks at a time); returns the output vectors."""
    out = []
    t_end = orbits * ORBIT
    for c0 in range(0, len(params), chunk):
        ps = params[c0:c0 + chunk]
        with tempfile.TemporaryDirectory() as tmp:
            t0 = time.perf_counter()
            with open_ensemble(ps, tmp) as ens:
                steps = ens.run(t_end)
                out.extend(to_output(ens, k, p) for k, p in enumerate(ps))
            if log:
                log({"generated": c0 + len(ps), "batch_substeps": steps, "seconds": round(time.perf_counter() - t0, 2)})
    return np.stack(out)


class Emulator:
    """The trained network plus the output normalisation."""

    def __init__(self, seed=7, max_batch=512):
        self.net = MLP(WIDTHS, max_batch=max_batch, seed=seed)
        self.mean = np.zeros(2 * NOUT, np.float32)
        self.std = np.ones(2 * NOUT, np.float32)

    def close(self):
        self.net.close()

    def fit(self, x, y, steps, lr=1e-3, batch=None):
        self.mean = y.mean(axis=0)
        self.std = y.std(axis=0) + 1e-6
        yn = ((y - self.mean) / self.std).astype(np.float32)
        batch = batch or x.shape[0]
        return self.net.train(x, yn, batch=batch, steps=steps, lr=lr)

    def predict(self, x):
        return self.net.forward(x) * self.std + self.mean

    def save(self, path):
        np.savez(path, params=self.net.params().astype(np.float32), mean=self.mean, std=self.std)

    def load(self, path):
        z = np.load(path)
        self.net.set_params(z["params"])
        self.mean, self.std = z["mean"], z["std"]


def warm_state(ens, k, p, y):
    """Initial state of disk k from the predicted output vector y."""
    dens0, vr0, vt0 = ens.state(k)
    R, ns = dens0.shape
    nr = R - 1
    rmed = ens.ring(k, "Rmed")[:nr]
    rinf = ens.ring(k, "Rinf")[:nr]
    sigmed = ens.ring(k, "SigmaMed")[:nr]
    ratio = np.maximum(1.0 + np.interp(np.log(rmed), np.log(R_OUT), y[:NOUT]), 1e-3)
    sig = sigmed * ratio
    vr = np.interp(np.log(rinf), np.log(R_OUT), y[NOUT:]) * vr_scale(p)
    vr[0] = 0.0
    # radial equilibrium: v_phi^2 / r = 1/r^2 + (dP/dr) / Sigma, P = c_s^2 Sigma,
    # c_s = h r^f v_K (locally isothermal)
    cs2 = (p["h"] * rmed ** p["flaring"]) ** 2 / rmed
    P = cs2 * sig
    dPdr = np.gradient(P, rmed)
    vphi = np.sqrt(np.maximum(rmed * (1.0 / rmed ** 2 + dPdr / sig), 0.0))
    omega_frame = ens.get(k, "OmegaFrame")
    dens = dens0.copy()
    vrad = vr0.copy()
    vth = vt0.copy()
    dens[:nr] = sig[:, None]
    vrad[:nr] = vr[:, None]
    vth[:nr] = (vphi - omega_frame * rmed)[:, None]
    return dens, vrad, vth


def orbit_history(ens, max_orbits, log=None, label=""):
    """Advance every disk max_orbits orbits; per orbit, the orbit average
    (over its 20 DT boundaries) of each disk's ring-averaged Sigma / Sigma_init
    and its steadiness against the previous orbit.  Returns R[disk][orbit],
    the orbit averages hist[disk][orbit - 1] and the wall time."""
    n = ens.n
    per_orbit = int(round(ORBIT / DT))
    rmed = [ens.ring(k, "Rmed")[:ens.shapes[k][0] - 1] for k in range(n)]
    t0 = time.perf_counter()
    t_start = max(ens.get(k, "PhysicalTime") for k in range(n))
    sigmed = [ens.ring(k, "SigmaMed")[:ens.shapes[k][0] - 1] for k in range(n)]
    prev = None
    R = [[] for _ in range(n)]
    hist = [[] for _ in range(n)]  # orbit averages, orbit 1..max_orbits
    avg = None
    for o in range(1, max_orbits + 1):
        acc = [0.0] * n
        for q in range(1, per_orbit + 1):
            ens.run(t_start + ((o - 1) * per_orbit + q) * DT)
            for k in range(n):
                acc[k] = acc[k] + ens.dens(k)[:-1].mean(axis=1) / sigmed[k]
        avg = [a / per_orbit for a in acc]
        for k in range(n):
            hist[k].append(avg[k])
        if prev is not None:
            for k in range(n):
                R[k].append(steadiness(prev[k], avg[k], rmed[k]))
        prev = avg
        if log and o % 25 == 0:
            log({"start": label, "orbit": o, "R_median": float(np.median([r[-1] for r in R])),
                 "seconds": round(time.perf_counter() - t0, 1)})
    return R, hist, time.perf_counter() - t0


def first_below(r, eps):
    """First orbit whose steadiness is below eps (orbit 2 is the first
    comparable one), or None."""
    for i, x in enumerate(r):
        if x < eps:
            return i + 2
    return None


def study(n_train, n_eval, nrad, orbits, steps, eps, max_orbits, seed, log=print):
    report = {"grid": [nrad, 3 * nrad], "n_train": n_train, "n_eval": n_eval, "train_orbits": orbits,
              "eps_per_orbit": list(eps), "max_orbits": max_orbits, "seed": seed, "widths": WIDTHS}
    train_p = draw(n_train, seed, nrad)
    eval_p = draw(n_eval, seed + 1, nrad)
    t0 = time.perf_counter()
    y_train = generate(train_p, orbits, log=log)
    report["generate_seconds"] = round(time.perf_counter() - t0, 1)
    x_train = features(train_p)
    emu = Emulator(max_batch=max(n_train, n_eval))
    t0 = time.perf_counter()
    losses = emu.fit(x_train, y_train, steps=steps)
    report["train_seconds"] = round(time.perf_counter() - t0, 2)
    report["loss_first_last"] = [float(losses[0]), float(losses[-1])]
    # held-out: profile error of the prediction vs the solver at `orbits`,
    # against the plain initial state as the prediction
    y_eval = generate(eval_p, orbits)
    pred = emu.predict(features(eval_p))
    err_emu = np.abs(pred[:, :NOUT] - y_eval[:, :NOUT]).max(axis=1)
    err_flat = np.abs(y_eval[:, :NOUT]).max(axis=1)
    report["sigma_err_at_train_orbits"] = {"emulator_median": float(np.median(err_emu)),
                                           "plain_ic_median": float(np.median(err_flat))}
    # steadiness histories from cold and warm starts (the warm start goes
    # through fg_sim_set_state, the restart path)
    import ctypes
    from .native import call
    _pd = ctypes.POINTER(ctypes.c_double)
    with tempfile.TemporaryDirectory() as tmp:
        with open_ensemble(eval_p, tmp) as cold:
            R_cold, f_cold, wall_cold = orbit_history(cold, max_orbits, log, "cold")
    with tempfile.TemporaryDirectory() as tmp:
        with open_ensemble(eval_p, tmp) as warm:
            for k, p in enumerate(eval_p):
                d, vr, vt = warm_state(warm, k, p, pred[k])
                call("fg_sim_set_state", warm.sim_handle(k), d.ctypes.data_as(_pd), vr.ctypes.data_as(_pd),
                     vt.ctypes.data_as(_pd), None)
            R_warm, f_warm, wall_warm = orbit_history(warm, max_orbits, log, "warm")
    report["wall_seconds_per_start"] = {"cold": round(wall_cold, 2), "warm": round(wall_warm, 2)}
    report["steadiness_cold"] = [[round(x, 7) for x in r] for r in R_cold]
    report["steadiness_warm"] = [[round(x, 7) for x in r] for r in R_warm]
    per_eps = {}
    for e in eps:
        tc = [first_below(r, e) for r in R_cold]
        tw = [first_below(r, e) for r in R_warm]
        both = [k for k in range(n_eval) if tc[k] and tw[k]]
        # profiles at the orbit where each start first meets the criterion
        steady_diff = []
        for k in both:
            pc, pw = f_cold[k][tc[k] - 1], f_warm[k][tw[k] - 1]
            steady_diff.append(float(np.abs(pw - pc).sum() / np.abs(pc).sum()))
        per_eps[str(e)] = {
            "steady_profile_l1_median": float(np.median(steady_diff)) if steady_diff else None,
            "steady_profile_within_1pct": int(sum(x <= 0.01 for x in steady_diff)),
            "reached_cold": int(sum(t is not None for t in tc)), "reached_warm": int(sum(t is not None for t in tw)),
            "orbits_cold": tc, "orbits_warm": tw,
            "median_speedup": float(np.median([tc[k] / tw[k] for k in both])) if both else None,
            "total_orbits_cold": int(sum(t if t else max_orbits for t in tc)),
            "total_orbits_warm": int(sum(t if t else max_orbits for t in tw)),
        }
    report["time_to_steady"] = per_eps
    diffs = []
    for k in range(n_eval):
        d = np.abs(f_warm[k][-1] - f_cold[k][-1])
        diffs.append({"l1": float(d.sum() / np.abs(f_cold[k][-1]).sum()),
                      "linf": float((d / np.abs(f_cold[k][-1])).max())})
    report["final_profile_diff"] = {"l1_median": float(np.median([x["l1"] for x in diffs])),
                                    "l1_max": float(max(x["l1"] for x in diffs)),
                                    "within_1pct_l1": int(sum(x["l1"] <= 0.01 for x in diffs)),
                                    "linf_median": float(np.median([x["linf"] for x in diffs]))}
    emu.close()
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description="emulator warm-start study")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("study")
    s.add_argument("--n-train", type=int, default=512)
    s.add_argument("--n-eval", type=int, default=32)
    s.add_argument("--nrad", type=int, default=128)
    s.add_argument("--orbits", type=float, default=100.0)
    s.add_argument("--steps", type=int, default=3000)
    s.add_argument("--eps", type=float, nargs="+", default=[3e-3, 2e-3, 1e-3, 5e-4, 2e-4])
    s.add_argument("--max-orbits", type=int, default=300)
    s.add_argument("--seed", type=int, default=20261202)
    s.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    def log(d):
        print(json.dumps(d), flush=True)

    rep = study(args.n_train, args.n_eval, args.nrad, args.orbits, args.steps, args.eps, args.max_orbits,
                args.seed, log)
    log(rep)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
