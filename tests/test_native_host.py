"""Host-side native code, built with g++ on the CPU (no GPU needed).

- The FP32 -> BF16 rounding the BF16 tiers use (sg_bf16.hpp), against
  torch's CPU conversion: round to nearest even, infinities kept, every
  NaN (including payloads only in the low 16 bits) stays a NaN.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "splinegrids.jl_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")


def _build(tmp_path, src, extra=()):
    exe = str(tmp_path / os.path.splitext(src)[0])
    cmd = ["g++", "-std=c++17", "-O1", "-g", *extra, "-I", CSRC,
           os.path.join(NATIVE, src), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_bf16_rounding_matches_torch(tmp_path):
    exe = _build(tmp_path, "bf16_check.cpp", ["-fsanitize=address,undefined",
                                              "-fno-sanitize-recover=all"])
    rng = np.random.default_rng(0)
    special = np.array([0, 0x80000000, 0x7f800000, 0xff800000, 0x7f7fffff, 0x7f7f8000,
                        0x3f808000, 0x3f818000, 0x3f807fff, 0x00000001, 0x807fffff,
                        0x7f800001, 0x7f80ffff, 0xff800001, 0x7fc00000, 0x7fffffff,
                        0xffc00001, 0x7f810000], dtype=np.uint64)
    bits = np.concatenate([special, rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)])
    out = subprocess.run([exe], input="\n".join(str(int(b)) for b in bits), check=True,
                         capture_output=True, text=True).stdout.split()
    got = np.array([int(x) for x in out], dtype=np.uint32)
    x = torch.from_numpy(bits.astype(np.uint32).view(np.float32))
    want = x.to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16).astype(np.uint32)
    nan = torch.isnan(x).numpy()
    got_f = torch.from_numpy((got << 16).astype(np.uint32).view(np.float32))
    assert bool(torch.isnan(got_f).numpy()[nan].all())  # NaN in, NaN out
    assert ((got[nan] >> 15) == (bits[nan].astype(np.uint32) >> 31)).all()  # sign kept
    assert np.array_equal(got[~nan], want[~nan])
