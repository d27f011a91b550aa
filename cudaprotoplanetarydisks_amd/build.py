"""Build libfargo_hip.so and the fargo_mi355x executable in-tree.

hipcc compiles the HIP sources for gfx950 only; g++ compiles the host
driver and the host tables.  Both use -ffp-contract=off so that no
multiply-add is fused:
the FP64 path rounds every operation like the reference's expression
order.  Run:  python -m cudaprotoplanetarydisks_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]
# cos and sin stay separate libm calls (no fused sincos, which can differ
# from sin in the last bit): the oracle is built the same way
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos", "-Wall",
             f"-I{INCLUDE}"]

HIP_SOURCES = ["fg_context.hip", "fg_steps.hip", "fg_reduce.hip", "fg_sg.hip", "fg_fused.hip", "fg_mlp.hip"]
HOST_SOURCES = ["fg_host.cpp", "fg_hostmath.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    so = os.path.join(LIB, "libfargo_hip.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, so)
    exe = os.path.join(LIB, "fargo_mi355x")
    _run(["g++"] + CXX_FLAGS + [os.path.join(CSRC, "fargo_main.cpp"), "-o", exe,
                                f"-L{LIB}", "-lfargo_hip", "-Wl,-rpath,$ORIGIN"])
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
