"""Build libsplinegrids_hip.so for gfx950 with hipcc.

Each translation unit under csrc/ is compiled in parallel into _build/, then
linked into a shared library next to this file.  Usable as a script
(``python splinegrids.jl_amd/build.py``) or through ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsplinegrids_hip.so")
# the write-bandwidth probe bench.py --workload ceiling runs (not in the library)
CEILING = os.path.join(HERE, "store_ceiling")
SOURCES = ["sg_basis.hip", "sg_forward.hip", "sg_mfma.hip", "sg_adjoint.hip", "sg_adjoint_mma.hip",
           "sg_batch.hip", "sg_rational.hip", "sg_noalloc.hip", "sg_refine.hip", "sg_plan_host.hip",
           "sg_api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (v_pk_* forms are not trusted here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: no v_accvgpr moves around each tile
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> tuple[str, str]:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source and link the library; returns its path."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}",
           *[o for o, _ in results], "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [HIPCC, *FLAGS, "-x", "hip", os.path.join(CSRC, "store_ceiling.hip"), "-o",
           CEILING + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for store_ceiling.hip:\n{r.stderr}")
    os.replace(CEILING + ".tmp", CEILING)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
