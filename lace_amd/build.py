"""Build recipe for liblace_mi355x.so (hipcc, gfx950 only).

Every source under csrc/ is compiled as HIP for gfx950 into _build/ and
linked into lace_amd/liblace_mi355x.so, which tests, smoke() and bench.py
load through ctypes.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblace_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    # keep the compiler from forming packed FP32 ops: some of their forms
    # have given silently wrong results on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-DLACE_BUILD",
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        # surface real warnings, drop the host pass's note about the feature flag
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print("\n".join(lines), file=sys.stderr)
    return obj


DRIVERS = os.path.join(HERE, "drivers")
DRIVER_BIN = os.path.join(DRIVERS, "bin")
# executable -> (source, extra defines): the reference's driver names
DRIVER_TARGETS = {
    "generate_iLU": ("generate_iLU.cpp", []),
    "generate_iLU_gpu": ("generate_iLU.cpp", ["-DLACE_DRIVER_GPU"]),
    "solver": ("solver_interface.cpp", []),
}


def build_drivers() -> list:
    """Host-only command-line drivers linked against the library (g++)."""
    os.makedirs(DRIVER_BIN, exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "include")
    out = []
    for exe, (src, defs) in DRIVER_TARGETS.items():
        target = os.path.join(DRIVER_BIN, exe)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", *defs, "-I", inc, os.path.join(DRIVERS, src),
               "-L", HERE, "-llace_mi355x", "-Wl,-rpath,$ORIGIN/../..", "-o", target]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(target)
    return out


def build(jobs: int = 8) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_drivers()
    return LIB


if __name__ == "__main__":
    print(build())
