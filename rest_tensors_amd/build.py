"""Builds librest_tensors_amd.so (HIP kernels for gfx950 + C ABI) in-tree.

Usage: python -m rest_tensors_amd.build  (or rest_tensors_amd.build.build()).
Every .hip / .cpp file under csrc/ is compiled with hipcc for gfx950 and
linked into rest_tensors_amd/librest_tensors_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librest_tensors_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
    f"--offload-arch={ARCH}",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Per-file flags.  The FP64 GEMM engines keep their 128 accumulator registers
# in VGPRs: with the default MFMA form the compiler holds them in AGPRs only
# around the MFMAs and copies all of them in and out every k-chunk.
FILE_FLAGS = {
    "gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
    "batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *FILE_FLAGS.get(os.path.basename(src), []), "-x", "hip", "-c",
           src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    build_host_mirror_test(verbose)
    return LIB


REPO = os.path.dirname(HERE)
MIRROR_TEST_SRC = os.path.join(REPO, "tests", "cpp", "test_host_mirror.cpp")
MIRROR_TEST_BIN = os.path.join(OBJ, "test_host_mirror")


def build_host_mirror_test(verbose: bool = True) -> str | None:
    """Compiles the C++ host-mirror test (include/rest_tensors.hpp) with g++
    against the library; the binary sits next to the objects."""
    if not os.path.exists(MIRROR_TEST_SRC):
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(REPO, "include"),
           MIRROR_TEST_SRC, "-L", HERE, "-lrest_tensors_amd",
           "-Wl,-rpath,$ORIGIN/..", "-o", MIRROR_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host mirror test build failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print(f"built {MIRROR_TEST_BIN}", flush=True)
    return MIRROR_TEST_BIN


if __name__ == "__main__":
    build()
    sys.exit(0)
