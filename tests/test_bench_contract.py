"""bench.py's output contract: one JSON line with the keys a caller reads.

The CPU test checks that a multi-GPU request is refused before any device
work; the GPU test runs a short default benchmark and checks the line.
"""
from __future__ import annotations

import contextlib
import io
import json

import pytest

import bench

KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
        "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config", "roofline"}
ROOFLINE = {"bound", "achieved", "peak", "unit", "frac", "traffic"}


def test_bench_refuses_several_gpus(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU only" in capsys.readouterr().err


def test_cpu_baseline_legs():
    """The CPU baseline reports the threads it used, the host (nproc, CPU
    model, cgroup quota), a one-core leg and an FP64 leg."""
    wl = bench.workload("c1")
    cb = bench.cpu_baseline(wl, budget_s=0.4)
    assert cb["kind"] == "port" and cb["unit"] == "sample-pts/s" and cb["value"] > 0
    threads, host = bench.host_cpus()
    assert cb["cores"] == threads >= 1
    assert {"nproc", "os_cpu_count", "cgroup_cpu_quota", "cpu_model"} <= set(cb["host"])
    assert cb["one_core"]["value"] > 0 and cb["fp64_all_cores"]["dtype"] == "float64"


@pytest.mark.gpu
def test_bench_default_line(gpu):
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rc = bench.main(["--steps", "5", "--warmup", "1", "--no-cpu-baseline"])
    assert rc == 0
    lines = [line for line in buf.getvalue().splitlines() if line.strip()]
    assert len(lines) == 1, lines
    d = json.loads(lines[0])
    assert KEYS <= set(d), KEYS - set(d)
    assert ROOFLINE <= set(d["roofline"])
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 1
    assert d["unit"] == "sample-pts/s" and d["higher_is_better"] is True
    assert d["dtype"] == "f32" and d["config"]["workload"].startswith("c3:")
    assert 0 < d["roofline"]["frac"] < 1
    assert d["value"] == pytest.approx(512 * 512 * 256 / (d["ms_per_step"] * 1e-3), rel=1e-9)
    assert max(d["max_rel_err_vs_fp64"]) < 1e-6
    assert {"native", "split_f16", "bf16"} <= set(d["tiers"])
