"""Builds the HIP engine (libmclicks.so) for gfx950 and its command-line
driver (mclicks) with hipcc.

The library is built in-tree (massiveclicks_amd/lib/) so that it travels with
the source tree; it is rebuilt on every call to :func:`build`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmclicks.so")
CLI = os.path.join(LIBDIR, "mclicks")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mc_engine.hip", "mc_eval.hip", "mc_grad.hip", "mc_layout.cpp", "mc_dataset.cpp", "mc_io.cpp", "mc_tune.cpp"]

# -ffp-contract=off keeps the E-step arithmetic in the reference's operation
# order; packed FP32 instructions are disabled on the device.
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {what}")


def build(verbose: bool = False) -> str:
    """Builds lib/libmclicks.so and the command-line driver lib/mclicks.
    Every source compiles to its own object, in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = max(1, min(len(SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda so: _run([HIPCC, *FLAGS, "-c", "-o", so[1], os.path.join(CSRC, so[0])], so[0], verbose),
                      zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + objs, "libmclicks.so", verbose)
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "mc_cli.cpp"),
          "-L", LIBDIR, "-lmclicks", "-Wl,-rpath,$ORIGIN", "-o", tmp], "mclicks", verbose)
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
