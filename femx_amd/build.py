"""Build femx_amd/libfemx_amd.so for gfx950 with hipcc (in-tree).

Every .hip file under csrc/ is compiled for gfx950 only and the host
builders (fx_host.cpp) for the host; all link into one shared library that
exports the C-ABI of include/femx_amd.h.  Floating-point contraction is off
unless a source asks for fused multiply-adds explicitly, and packed FP32
instructions are disabled.

    python -m femx_amd.build        # or femx_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfemx_amd.so")
OBJ = os.path.join(HERE, "_build")

ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-x", "hip",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc(), "-c", src, "-o", obj] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_cpp_tests()
    return OUT


CPP_TEST = os.path.join(OBJ, "test_femx_hip")


def build_cpp_tests():
    """The C++ interface tests (tests/cpp), linked against the library."""
    src = os.path.join(ROOT, "tests", "cpp", "test_femx_hip.cpp")
    cmd = [hipcc(), src, "-o", CPP_TEST, "-O2", "-std=c++17", "--offload-arch=" + ARCH, "-I" + os.path.join(HERE, "include"),
           "-I" + os.path.join(ROOT, "include"), "-L" + HERE, "-lfemx_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CPP_TEST


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
