"""The host plan code and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU (no GPU).

- tests/native/plan_check.cpp drives splinegrids.jl_amd/csrc/sg_plan_host.hip
  (tilings, LDS budgets, the adjoint's gather lists, the transposed band
  lists, the batch's tile windows and queue packing) over the shapes the GPU
  tests use plus corner cases, and checks the bounds the kernels rely on.
  Built with hipcc, the sanitizers on the host code only.
- tests/native/oracle_check.c runs every entry of oracle/sg_oracle.c and
  checks it against what the mathematics fixes (partition of unity, the
  constant spline, transpose identities, the Flag scatter), with leak
  detection on.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "splinegrids.jl_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")
ORACLE = os.path.join(ROOT, "oracle")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(exe, env_extra):
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_plan_host_code_under_sanitizers(tmp_path):
    if not (shutil.which(HIPCC) or os.path.exists(HIPCC)):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "plan_check")
    cmd = [HIPCC, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all",
           "-x", "hip", "-I", CSRC, os.path.join(CSRC, "sg_plan_host.hip"),
           os.path.join(NATIVE, "plan_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
    # the HIP runtime library keeps allocations to the end of the process
    out = _run(exe, {"ASAN_OPTIONS": "detect_leaks=0", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert out.startswith("plan_check ok"), out


def test_oracle_under_sanitizers(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("gcc not found")
    exe = str(tmp_path / "oracle_check")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", ORACLE, os.path.join(ORACLE, "sg_oracle.c"),
           os.path.join(NATIVE, "oracle_check.c"), "-o", exe, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    out = _run(exe, {"ASAN_OPTIONS": "detect_leaks=1", "OMP_NUM_THREADS": "4"})
    assert out.startswith("oracle_check ok"), out
