"""Build libkhs.so (HIP, gfx950) and the khs_run driver in-tree with hipcc.

    python kh-simulations_amd/build.py [--jobs N]

Objects go to kh-simulations_amd/_build/, the library and the driver to
kh-simulations_amd/khs/ next to the ctypes binding that loads them.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
DRIVER = os.path.join(HERE, "driver")
OUT = os.path.join(HERE, "khs")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
# Device code: gfx950 only; packed-FP32 instructions disabled.
HIP = ["-x", "hip", f"--offload-arch={ARCH}",
       "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("khs_step_kernels.hip", True),
    ("khs_aux_kernels.hip", True),
    ("khs_filter_mfma.hip", True),
    ("khs_context.hip", True),
    ("khs_config.cpp", False),
    ("khs_init.cpp", False),
    ("khs_analysis.cpp", False),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n  " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def compile_one(src, is_hip):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + (HIP if is_hip else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build(jobs=4, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(*s), SOURCES))
    lib = os.path.join(OUT, "libkhs.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, lib)
    exe = os.path.join(OUT, "khs_run")
    _run([HIPCC, "-O2", "-std=c++17", os.path.join(DRIVER, "khs_run.cpp"), "-o", exe,
          "-L" + OUT, "-lkhs", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {lib}\nbuilt {exe}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
