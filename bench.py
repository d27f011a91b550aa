"""Benchmark of the FARGO substep on one MI355X.

A step is one substep (one iteration of AlgoGas's CFL sub-cycle,
SourceEuler.cu:222) over the whole grid.  Workload: the precision-mode
grid, 1024 x 3072 (r in [0.4, 2.5], log spacing) with the plain-FARGO
physics (isothermal h = 0.05, nu = 1e-5, one q = 1e-3 planet, evanescent
edges) and a seeded 1e-4 white-noise perturbation of Sigma.  Inputs are
resident in HBM when the timed region starts.

Prints one JSON line: zone-updates/s of the GPU path (FP64 by default),
the roofline of the dominant kernel (HIP-event timing inside the library),
the oracle's multicore FP64 rate on a bounded sample of the same workload,
the same substep in the other precision ("other_precision": the FP32 bulk
mode beside the FP64 headline) and a batched ensemble of drawn disks.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6  # MI355X dense FP64 vector
SEED = 20261202
METRIC = "zone-updates/s + % HBM peak; ensemble time-to-steady-gap, emulator vs plain"
KERNELS = ["k_src", "k_trad", "k_taz", "k_fin", "force_reduce", "cfl_count"]


def algorithmic_bytes(nr, ns, precision):
    """Compulsory HBM bytes of one launch of each fused kernel: every
    input field read once and every output field written once, over the
    (NRAD+1) x NSEC field.  The state between substeps is FP64 in both
    modes; the FP32 mode stores the fields passed between the kernels
    (after k_src, k_trad and k_taz) in FP32."""
    S = (nr + 1) * ns
    if precision == "fp32":
        return {
            "k_src": S * (3 * 8 + 3 * 4),   # FP64 state in; Sigma, vrad, vtheta out (FP32)
            "k_trad": S * (3 * 4 + 5 * 4),  # Sigma, vrad, vtheta in; Sigma + 4 momenta out
            "k_taz": S * (6 * 4 + 4 * 4),   # 5 quantities + vtheta in; 4 out
            "k_fin": S * (4 * 4 + 3 * 8),   # 4 in; final FP64 state out
            "force_reduce": nr * ns * 8,
            "cfl_count": nr * 8 * 2,
        }
    return {
        "k_src": S * 8 * (3 + 3),   # Sigma, vrad, vtheta in and out
        "k_trad": S * 8 * (3 + 5),  # state in; Sigma + 4 momenta out
        "k_taz": S * 8 * (6 + 4),   # 5 quantities + vtheta in; 4 out
        "k_fin": S * 8 * (4 + 3),   # 4 in; final state out
        "force_reduce": nr * ns * 8,
        "cfl_count": nr * 8 * 2,
    }


def write_workload_deck(tmp, nr, ns):
    from cudaprotoplanetarydisks_amd.workloads import write_deck
    return write_deck(os.path.join(tmp, "workload.par"), Nrad=nr, Nsec=ns, OutputDir=tmp + "/")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(deck, nr, ns, steps, warm, omp=True):
    """Oracle (FP64, same arithmetic; OpenMP per-cell loops, or the serial
    build) on the same noisy initial state."""
    import numpy as np
    from oracle_binding import Oracle
    orc = Oracle(deck, omp=omp)
    dens = orc.view("Dens")
    sig = orc.field("SigmaMed")
    idx = (np.arange(nr, dtype=np.uint64)[:, None] * np.uint64(ns) + np.arange(ns, dtype=np.uint64)[None, :])
    x = np.uint64(SEED) ^ idx
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    xi = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
    dens[:nr] = sig[:nr, None] * (1.0 + 1e-4 * xi)
    for _ in range(warm):
        orc.substep()
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.substep()
    t1 = time.perf_counter()
    orc.close()
    return nr * ns * steps / (t1 - t0), t1 - t0


def ensemble_leg(disks, steps, warmup, seed=SEED):
    """Batched ensemble of drawn disks (uneven grids 128..384 x 3 NRAD, FP32
    bulk): aggregate zone-updates/s over `steps` batch substeps, all disks
    active (the target time is far)."""
    from cudaprotoplanetarydisks_amd.ensemble import make_ensemble
    ens, params, _ = make_ensemble(disks, seed)
    try:
        t_end = 1e30
        for _ in range(warmup):
            ens.substep(t_end)
        sub0 = [ens.get(k, "substeps") for k in range(ens.n)]
        t0 = time.perf_counter()
        for _ in range(steps):
            ens.substep(t_end)
        wall = time.perf_counter() - t0
        zones = [(R - 1) * ns for R, ns in ens.shapes]
        zu = sum(z * (ens.get(k, "substeps") - s0) for k, (z, s0) in enumerate(zip(zones, sub0)))
        grids = sorted({p["nrad"] for p in params})
        return {"disks": disks, "seed": seed, "grids_nrad": grids, "zones": int(sum(zones)),
                "batch_substeps": steps, "ms_per_batch_substep": round(1e3 * wall / steps, 4),
                "value": round(zu / wall, 1), "unit": "zone-updates/s", "dtype": "f32",
                "launch": "one captured hipGraph per batch substep, one launch per phase over the (disk, tile) work list"}
    finally:
        ens.close()


def gpu_leg(deck, precision, path, steps, warmup):
    """Time `steps` substeps after `warmup` (inputs resident in HBM), then a
    separate pass with per-kernel HIP events.  Returns (zone-updates/s,
    seconds, {kernel: ms per launch}, all-finite)."""
    import numpy as np
    from cudaprotoplanetarydisks_amd import Simulation, lib
    from cudaprotoplanetarydisks_amd.native import call
    sim = Simulation(deck, fused=(path == "fused"), precision=precision, noise_seed=SEED, noise_amp=1e-4)
    try:
        ctx = sim.ctx.h
        for _ in range(warmup):
            sim.substep()
        call("fg_sync", ctx)
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.substep()
        call("fg_sync", ctx)
        elapsed = time.perf_counter() - t0
        # kernel timing (separate pass so events do not perturb the clock)
        prof = {}
        if path == "fused":
            call("fg_profile", ctx, 1)
            for _ in range(min(steps, 50)):
                sim.substep()
            ms = (ctypes.c_double * 6)()
            calls = (ctypes.c_long * 6)()
            n = lib().fg_profile_read(ctx, ms, calls, 6)
            call("fg_profile", ctx, 0)
            prof = {KERNELS[k]: ms[k] / max(1, calls[k]) for k in range(n)}
        d, _, _ = sim.state()
        nr, ns = sim.nrad, sim.nsec
        return nr * ns * steps / elapsed, elapsed, prof, bool(np.isfinite(d).all())
    finally:
        sim.close()


def deck_copy(tmp, name):
    """A committed deck with its planet file and output directory made
    absolute (the decks name them relative to the decks directory)."""
    from cudaprotoplanetarydisks_amd.workloads import DECKS
    src = open(os.path.join(DECKS, name)).read()
    for cfg in ("jupiter_feeldisk.cfg", "jupiter_fixed.cfg"):
        src = src.replace(cfg, os.path.join(DECKS, cfg))
    src = src.replace("out_c1/", tmp + "/").replace("out_c2/", tmp + "/")
    deck = os.path.join(tmp, name)
    with open(deck, "w") as f:
        f.write(src)
    return deck


def small_grid_leg(steps, warmup):
    """The plain-FARGO workload at its own 128 x 384 grid (decks/c1_plain_fargo.par)
    on the fused FP64 path: at this size the per-substep host round trip
    (CFL count and star force back to the host, planets advanced there,
    the next substep launched) is part of the time."""
    from cudaprotoplanetarydisks_amd import Simulation
    from cudaprotoplanetarydisks_amd.native import call
    with tempfile.TemporaryDirectory() as tmp:
        with Simulation(deck_copy(tmp, "c1_plain_fargo.par"), fused=True) as sim:
            ctx = sim.ctx.h
            for _ in range(warmup):
                sim.substep()
            call("fg_sync", ctx)
            t0 = time.perf_counter()
            for _ in range(steps):
                sim.substep()
            call("fg_sync", ctx)
            el = time.perf_counter() - t0
            zones = sim.nrad * sim.nsec
        return {"workload": "plain FARGO, 128x384 (decks/c1_plain_fargo.par), fused FP64 path",
                "steps": steps, "ms_per_step": round(1e3 * el / steps, 4),
                "value": round(zones * steps / el, 1), "unit": "zone-updates/s"}


def adsg_leg(steps, warmup):
    """The ADSG workload (decks/c2_adsg.par: 512 x 1536, adiabatic with beta
    cooling, alpha viscosity, self-gravity on the 1024 x 1536 padded grid,
    a planet feeling the disk) on the fused FP64 path, and on the
    per-routine path for comparison: zone-updates/s and the share of HBM
    peak at the compulsory 64 B/zone (Sigma, v_r, v_theta, e read and
    written) + 576 B/zone of self-gravity FFT traffic."""
    from cudaprotoplanetarydisks_amd import Simulation
    from cudaprotoplanetarydisks_amd.native import call

    def run(fused):
        with tempfile.TemporaryDirectory() as tmp:
            with Simulation(deck_copy(tmp, "c2_adsg.par"), fused=fused) as sim:
                assert sim.get("fused_active") == (1.0 if fused else 0.0)
                ctx = sim.ctx.h
                for _ in range(warmup):
                    sim.substep()
                call("fg_sync", ctx)
                t0 = time.perf_counter()
                for _ in range(steps):
                    sim.substep()
                call("fg_sync", ctx)
                el = time.perf_counter() - t0
                return sim.nrad * sim.nsec * steps / el, el

    zus, el = run(True)
    rzus, rel = run(False)
    return {"workload": "ADSG + self-gravity, 512x1536 (decks/c2_adsg.par), fused FP64 path",
            "steps": steps, "ms_per_step": round(1e3 * el / steps, 4), "value": round(zus, 1),
            "unit": "zone-updates/s", "b_alg_per_zone": 64 + 576,
            "pct_hbm_peak": round(100.0 * zus * (64 + 576) / (HBM_PEAK_GBS * 1e9), 3),
            "per_routine": {"ms_per_step": round(1e3 * rel / steps, 4), "value": round(rzus, 1)}}


def emulator_leg(n_eval, max_orbits):
    """Ensemble time-to-steady, emulator warm start against plain cold start:
    held-out draws (128 x 384 disks, the training seed + 1) evolved by the
    batched solver from both starts with the committed weights
    (cudaprotoplanetarydisks_amd/data/emulator_128x384.npz, trained on an
    MI355X by the project's BF16-MFMA kernels on 512 batched runs of 1000
    orbits).  Steady: the mass-weighted L1 change of the orbit-averaged ring
    profile per orbit below eps."""
    from cudaprotoplanetarydisks_amd.emulator import WEIGHTS, warm_start_leg
    t0 = time.perf_counter()
    rep = warm_start_leg(WEIGHTS, n_eval, max_orbits, [1e-4, 5e-5, 1e-5])
    rep["seconds"] = round(time.perf_counter() - t0, 1)
    return rep


def roofline_of(prof, nr, ns, precision):
    """Roofline entry of the dominant (longest) kernel: algorithmic bytes
    per launch (algorithmic_bytes) over its event-timed duration."""
    if not prof:
        return None
    dom = max(prof, key=prof.get)
    ab = algorithmic_bytes(nr, ns, precision)[dom]
    achieved = ab / (prof[dom] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(f"{precision}:{dom}")
    out = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "kernel_ms": {k: round(v, 4) for k, v in prof.items()}}
    # the compute side of the same launch: executed FP64 operations per launch
    # (rocprofv3 SQ_INSTS_VALU_*_F64 counters at 1024 x 3072, committed in
    # profiles/fp64_ops.json) over the live event-timed duration
    ofile = os.path.join(ROOT, "profiles", "fp64_ops.json")
    if precision == "fp64" and (nr, ns) == (1024, 3072) and os.path.exists(ofile):
        with open(ofile) as f:
            ops = json.load(f).get(f"fp64:{dom}")
        if ops:
            tf = ops / (prof[dom] * 1e-3) / 1e12
            out["fp64"] = {"ops_per_launch": ops, "achieved": round(tf, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                           "frac": round(tf / FP64_PEAK_TFS, 4)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nrad", type=int, default=1024)
    ap.add_argument("--nsec", type=int, default=3072)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--path", choices=["fused", "routine"], default="fused")
    ap.add_argument("--cpu-steps", type=int, default=160)
    ap.add_argument("--cpu-serial-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-alt", action="store_true", help="skip the leg in the other precision")
    ap.add_argument("--ensemble-disks", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--adsg-steps", type=int, default=50, help="substeps of the ADSG leg (0: skip)")
    ap.add_argument("--small-steps", type=int, default=1000, help="substeps of the 128x384 leg (0: skip)")
    ap.add_argument("--emulator-disks", type=int, default=32, help="held-out disks of the warm-start leg (0: skip)")
    ap.add_argument("--emulator-orbits", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(every GPU box here has one GPU)", file=sys.stderr)
        sys.exit(2)
    nr, ns = args.nrad, args.nsec
    with tempfile.TemporaryDirectory() as tmp:
        deck = write_workload_deck(tmp, nr, ns)
        zus, elapsed, prof, finite = gpu_leg(deck, args.precision, args.path, args.steps, args.warmup)
        roofline = roofline_of(prof, nr, ns, args.precision)
        # the other precision on the same workload, reported beside the headline
        alt = None
        if not args.no_alt and args.path == "fused":
            other = "fp32" if args.precision == "fp64" else "fp64"
            a_zus, a_el, a_prof, a_fin = gpu_leg(deck, other, "fused", args.steps, args.warmup)
            alt = {"precision": other, "value": round(a_zus, 1), "unit": "zone-updates/s",
                   "ms_per_step": round(1e3 * a_el / args.steps, 4), "finite": a_fin,
                   "roofline": roofline_of(a_prof, nr, ns, other)}
        ens = None
        if args.ensemble_disks > 0:
            ens = ensemble_leg(args.ensemble_disks, args.ensemble_steps, 10)
        adsg = adsg_leg(args.adsg_steps, 2) if args.adsg_steps > 0 else None
        small = small_grid_leg(args.small_steps, 20) if args.small_steps > 0 else None
        emu = emulator_leg(args.emulator_disks, args.emulator_orbits) if args.emulator_disks > 0 else None
        cpu = None
        if not args.no_cpu:
            threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
            os.environ["OMP_NUM_THREADS"] = str(threads)
            rate, secs = cpu_baseline(deck, nr, ns, args.cpu_steps, 2)
            cpu = {"value": round(rate, 1), "unit": "zone-updates/s", "cores": threads, "kind": "port",
                   "sample": f"oracle FP64 (OpenMP), {nr}x{ns}, 2 warm-up + {args.cpu_steps} substeps "
                             f"({secs:.1f} s)", "cpu_model": cpu_model()}
            if args.cpu_serial_steps > 0:
                srate, ssecs = cpu_baseline(deck, nr, ns, args.cpu_serial_steps, 2, omp=False)
                cpu["serial"] = {"value": round(srate, 1), "unit": "zone-updates/s", "cores": 1,
                                 "sample": f"serial oracle, {nr}x{ns}, 2 warm-up + {args.cpu_serial_steps} "
                                           f"substeps ({ssecs:.1f} s)"}
    b_alg = 48  # read + write of Sigma, vrad, vtheta (FP64) per zone-update
    out = {
        "metric": METRIC, "value": round(zus, 1), "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": f"plain FARGO substep, {nr}x{ns} log grid, isothermal h=0.05, nu=1e-5, "
                               "q=1e-3 planet, evanescent edges, seeded 1e-4 noise on Sigma",
                   "nrad": nr, "nsec": ns, "path": args.path, "precision": args.precision,
                   "state": "fp64"},
        "pct_hbm_peak": round(100.0 * zus * b_alg / (HBM_PEAK_GBS * 1e9), 3),
        "roofline": roofline, "cpu_baseline": cpu, "ensemble": ens, "adsg": adsg, "small_grid": small,
        "emulator": emu,
        "gpu_vs_cpu": round(zus / cpu["value"], 1) if cpu else None,
        "finite": finite,
        "other_precision": alt,
    }
    if alt and cpu:
        alt["gpu_vs_cpu"] = round(alt["value"] / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
