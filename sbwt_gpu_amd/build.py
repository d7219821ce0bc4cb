"""Builds libsbwt_gpu_amd.so for gfx950 from csrc/*.hip with hipcc.

    python -m sbwt_gpu_amd.build [--jobs N]

Objects go to sbwt_gpu_amd/_build/, the library to sbwt_gpu_amd/ (in-tree,
so it travels with the source tree). hipcc cross-compiles for gfx950 without
a GPU present.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsbwt_gpu_amd.so")
ARCH = os.environ.get("SBWT_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]

# Per-file extras. mlp.hip: matrix-core accumulators in VGPRs (the vector
# ALU reads them every layer; AGPR accumulators would cost one
# v_accvgpr_read per element).
EXTRA = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + [
        os.path.join(os.path.dirname(HERE), "include", "sbwt_gpu_amd.h")]
    cmd = [HIPCC, *CFLAGS, *EXTRA.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(jobs: int = 0, verbose: bool = False) -> str:
    """Compiles every source and links the library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    # the command-line search over the C++ host API (include/sbwt_gpu_amd.hpp)
    cli = os.path.join(HERE, "sbwt_search")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "host", "sbwt_search.cpp"), "-o", cli + ".tmp", "-L", HERE, "-lsbwt_gpu_amd",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sbwt_search build failed:\n{r.stderr[-6000:]}")
    os.replace(cli + ".tmp", cli)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
