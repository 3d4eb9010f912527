"""Builds libcgsense_hip.so for gfx950 with hipcc (run by __graft_entry__.build()).

The library is built in-tree (codeare_amd/lib/) so that it travels with the
repository snapshot; it is compiled every time build() runs.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libcgsense_hip.so")
SOURCES = ["plan.hip", "kernels_fft.hip", "kernels_fft_reg.hip", "kernels_toep.hip", "kernels_grid.hip", "kernels_vec.hip",
           "kernels_sweep.hip", "kernels_b0.hip", "kernels_bin.hip",
           "unet.hip"]
ARCH = os.environ.get("CGS_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()

    def compile_one(src: str) -> str:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_cpp_test() -> str:
    """Builds tests/cpp/t_ncsense (the C++ operator facade driven like codeare's
    CGSENSE module) against the in-tree library; run after build()."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "t_ncsense.cpp")
    out = os.path.join(root, "tests", "cpp", "t_ncsense")
    gxx = shutil.which("g++") or "g++"
    cmd = [gxx, "-O2", "-std=c++17", "-Wall", src, "-I", os.path.join(root, "include"),
           "-L", OUT_DIR, "-lcgsense_hip", "-Wl,-rpath,$ORIGIN/../../codeare_amd/lib",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on t_ncsense.cpp:\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_test())
