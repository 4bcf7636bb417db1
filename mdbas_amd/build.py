"""Builds libmdgpu.so (the HIP kernels and the C ABI) for gfx950, in-tree.

Each ``csrc/*.hip`` file is compiled to an object with hipcc in parallel,
the host-only ``csrc/*.cpp`` files (reference-format I/O) with g++, and
all are linked against rocFFT.  All HIP code is built with packed-FP32
instructions disabled (``-target-feature -packed-fp32-ops``).  The C host
driver ``host/mdbas_gpu.c`` is then built with gcc against the library.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmdgpu.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("MDG_ARCH", "gfx950")

CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wno-unused-result", "-I" + os.path.join(ROCM, "include")]


HOSTFLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]
DRIVER_SRC = os.path.join(HERE, "host", "mdbas_gpu.c")
DRIVER = os.path.join(HERE, "mdbas_gpu")


def _compile(src: str) -> str:
    base, ext = os.path.splitext(os.path.basename(src))
    obj = os.path.join(OBJ, base + (".o" if ext == ".hip" else ".host.o"))
    if ext == ".cpp":
        cmd = ["g++", *HOSTFLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs,
           "-L" + os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if os.path.exists(DRIVER_SRC):
        cmd = ["gcc", "-std=gnu99", "-O2", "-Wall", "-o", DRIVER + ".tmp", DRIVER_SRC,
               "-I" + os.path.join(HERE, "..", "include"), "-L" + HERE, "-lmdgpu", "-lpthread",
               "-lm", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host driver build failed:\n{r.stderr[-4000:]}")
        os.replace(DRIVER + ".tmp", DRIVER)
        if verbose:
            print("built", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
