"""Benchmark of the grid-evaluation hot path on one MI355X.

Default workload (one "step" = one pass of the hot path over one batch):
the 3-D -> 3 NURBS grid, 64 x 64 x 32 control points, degree (3, 3, 2),
512 x 512 x 256 samples, clamped uniform knots, weights U[0.5, 2), control
points U[-1, 1) in FP32, one evaluate! (derivative order 0) with the
contraction in plain FP32 (the headline value), then the split-FP16 (hi+lo)
and BF16 MFMA tiers timed the same way on the same inputs, each with its
max and RMS error against this library's FP64 path.  value = sample-points/s
= samples / time.  Inputs are resident in HBM before timing starts.

Prints one JSON line.  Other workloads (--workload c1|c2|c3n|c3d|c3a|c4|c4a|c5)
are for profiling and parity runs; copy|fill calibrate the HBM counters and
ceiling prints the write bandwidth of the forward's bytes as a plain fill and
in the forward kernel's store pattern.
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0
METRIC = "sample-pts/s + HBM GB/s %peak; split-FP16 MFMA vs FP32 at ≤1e-6 rel err"
SEED0 = 0x5EED0000


# --------------------------------------------------------------------------
# workloads (shared with tests/test_gpu_configs.py)
# --------------------------------------------------------------------------

def c2_knot_values(rng, n=128, p=3, n_doubled=12):
    """Clamped knots: end multiplicity p+1, sorted U[0,1) interior values of
    which n_doubled have multiplicity 2 (132 knots for n = 128, p = 3)."""
    n_inner = n - p - 1 - n_doubled  # interior slots minus the extra copies
    while True:
        inner = np.sort(rng.uniform(0.0, 1.0, n_inner)).astype(np.float32)
        vals = np.concatenate([[0.0], inner, [1.0]]).astype(np.float32)
        if len(np.unique(vals)) == len(vals):
            break
    mults = np.ones(len(vals), dtype=np.int32)
    mults[0] = mults[-1] = p + 1
    mults[1 + rng.choice(n_inner, n_doubled, replace=False)] = 2
    assert mults.sum() == n + p + 1
    return vals, mults


def workload(name: str):
    """Description of a named workload: dims, arrays, fields, dtype."""
    if name == "c1":  # README shape, FP64, no derivatives
        rng = np.random.default_rng(SEED0 + 1)
        return dict(name="c1: 3-D->4 B-spline, n=(10,10,5), p=(2,3,2), m=(50,50,25), FP64",
                    n=(10, 10, 5), p=(2, 3, 2), m=(50, 50, 25), D=(0, 0, 0), nout=4,
                    dtype=np.float64, knots=None, fields=[(0, 0, 0)],
                    P=rng.uniform(0, 1, (10, 10, 5, 4)), W=None)
    if name == "c2":
        rng = np.random.default_rng(SEED0 + 2)
        knots = [c2_knot_values(rng), c2_knot_values(rng)]
        P = rng.standard_normal((128, 128, 1)).astype(np.float32)
        return dict(name="c2: 2-D->1 freeform surface, n=(128,128), p=(3,3), m=(4096,4096), "
                         "4 derivative fields, FP32",
                    n=(128, 128), p=(3, 3), m=(4096, 4096), D=(1, 1), nout=1,
                    dtype=np.float32, knots=knots,
                    fields=[(0, 0), (1, 0), (0, 1), (1, 1)], P=P, W=None)
    if name in ("c3", "c3n"):
        rng = np.random.default_rng(SEED0 + 3)
        P = rng.uniform(-1, 1, (64, 64, 32, 3)).astype(np.float32)
        W = rng.uniform(0.5, 2.0, (64, 64, 32)).astype(np.float32)
        return dict(name="c3: 3-D->3 NURBS, n=(64,64,32), p=(3,3,2), m=(512,512,256), FP32",
                    n=(64, 64, 32), p=(3, 3, 2), m=(512, 512, 256), D=(0, 0, 0), nout=3,
                    dtype=np.float32, knots=None, fields=[(0, 0, 0)], P=P,
                    W=W if name == "c3" else None)
    if name == "c3a":  # adjoint of the NURBS map in the control points (an extension)
        rng = np.random.default_rng(SEED0 + 3)
        P = rng.uniform(-1, 1, (64, 64, 32, 3)).astype(np.float32)
        W = rng.uniform(0.5, 2.0, (64, 64, 32)).astype(np.float32)
        return dict(name="c3a: adjoint of the 3-D->3 NURBS map (weights fixed), n=(64,64,32), "
                         "p=(3,3,2), m=(512,512,256), FP32",
                    n=(64, 64, 32), p=(3, 3, 2), m=(512, 512, 256), D=(0, 0, 0), nout=3,
                    dtype=np.float32, knots=None, fields=[(0, 0, 0)], P=P, W=W, adjoint=True)
    if name == "c3d":  # gradient of the NURBS map (quotient rule, an extension)
        rng = np.random.default_rng(SEED0 + 3)
        P = rng.uniform(-1, 1, (64, 64, 32, 3)).astype(np.float32)
        W = rng.uniform(0.5, 2.0, (64, 64, 32)).astype(np.float32)
        return dict(name="c3d: gradient of the 3-D->3 NURBS map, n=(64,64,32), p=(3,3,2), "
                         "m=(512,512,256), fields (1,0,0), (0,1,0), (0,0,1), FP32",
                    n=(64, 64, 32), p=(3, 3, 2), m=(512, 512, 256), D=(1, 1, 1), nout=3,
                    dtype=np.float32, knots=None,
                    fields=[(1, 0, 0), (0, 1, 0), (0, 0, 1)], P=P, W=W, rational=True)
    if name in ("copy", "fill"):  # calibration: plain device copy / fill of 1 GiB
        what = "device-to-device copy" if name == "copy" else "fill (write only)"
        return dict(name=f"{name}: 1 GiB {what} (bandwidth and counter calibration)",
                    n=(1,), p=(0,), m=(1 << 28,), D=(0,), nout=1, dtype=np.float32,
                    knots=None, fields=[(0,)], P=None, W=None, copy=name)
    if name == "c5":
        return dict(name="c5: 1024 uneven grids (half 2-D, half 3-D, 25 % NURBS, Nout 1|3), "
                         "one persistent launch, hipGraph, FP32",
                    batch=True, dtype=np.float32, fields=[None], W=None)
    if name in ("c4", "c4a"):
        rng = np.random.default_rng(SEED0 + 4)
        P = rng.uniform(-1, 1, (40, 40, 20, 3)).astype(np.float32)
        return dict(name="c4: 3-D->3 B-spline, n=(40,40,20), p=3, m=(400,400,200), FP32"
                         + (", adjoint" if name == "c4a" else ""),
                    n=(40, 40, 20), p=(3, 3, 3), m=(400, 400, 200), D=(0, 0, 0), nout=3,
                    dtype=np.float32, knots=None, fields=[(0, 0, 0)], P=P, W=None,
                    adjoint=(name == "c4a"))
    raise ValueError(f"unknown workload {name}")


def c5_specs(rng, n_grids=1024, small=False):
    """The uneven batch: half 2-D, half 3-D grids; control points per dimension
    U{8..64}; degree U{1..4} capped at n-1; samples per dimension U{32..512}
    (2-D) or U{16..128} (3-D) (smaller ranges with small=True, for tests);
    Nout in {1, 3}; 25 % NURBS."""
    specs = []
    for i in range(n_grids):
        nin = 2 if i < n_grids // 2 else 3
        n = tuple(int(x) for x in rng.integers(8, 65, nin))
        p = tuple(int(min(rng.integers(1, 5), nd - 1)) for nd in n)
        lo, hi = ((32, 513) if nin == 2 else (16, 129)) if not small else \
            ((16, 81) if nin == 2 else (8, 25))
        m = tuple(int(x) for x in rng.integers(lo, hi, nin))
        specs.append(dict(n=n, p=p, m=m, nout=int(rng.choice([1, 3])),
                          nurbs=bool(rng.random() < 0.25)))
    return specs


def build_spec_grid(sg, spec, rng, dtype=np.float32, device="cuda"):
    """A grid of one batch spec with P ~ U[-1, 1) and W ~ U[0.5, 2]."""
    dims = []
    for nd, pd, md in zip(spec["n"], spec["p"], spec["m"]):
        kw = {}
        if spec.get("unsorted"):
            kw["sample_points"] = rng.permutation(np.linspace(0.0, 1.0, md))
        dims.append(sg.SplineDimension(nd, pd, md, float_type=dtype, device=device, **kw))
    grid = (sg.NURBSGrid if spec["nurbs"] else sg.SplineGrid)(tuple(dims), spec["nout"])
    grid.control_points.copy_(rng.uniform(-1, 1, tuple(spec["n"]) + (spec["nout"],)).astype(dtype))
    if spec["nurbs"]:
        grid.weights.copy_(sg.from_numpy_colmajor(
            rng.uniform(0.5, 2.0, spec["n"]).astype(dtype), grid.dtype, device))
    return grid


def build_grid(sg, wl, device="cuda"):
    import torch

    dims = []
    for i in range(len(wl["n"])):
        kw = {}
        if wl["knots"] is not None:
            vals, mults = wl["knots"][i]
            kw["knot_vector"] = sg.KnotVector(vals.astype(wl["dtype"]), mults)
        dims.append(sg.SplineDimension(wl["n"][i], wl["p"][i], wl["m"][i],
                                       max_derivative_order=wl["D"][i],
                                       float_type=wl["dtype"], device=device, **kw))
    grid = (sg.NURBSGrid if wl["W"] is not None else sg.SplineGrid)(tuple(dims), wl["nout"])
    grid.control_points.copy_(wl["P"].astype(wl["dtype"]))
    if wl["W"] is not None:
        grid.weights.copy_(sg.from_numpy_colmajor(wl["W"], grid.dtype, device))
    torch.cuda.synchronize()
    return grid


def algorithmic_bytes(wl, adjoint=False) -> int:
    """Bytes one pass must move at minimum: every output written once, every
    input read once, per-dimension spans and value-type tables read once."""
    s = np.dtype(wl["dtype"]).itemsize
    nsamp, ncp = math.prod(wl["m"]), math.prod(wl["n"])
    F, nout = len(wl["fields"]), wl["nout"]
    tables = sum(m * (p + 1) * (D + 1) * s + 4 * m
                 for m, p, D in zip(wl["m"], wl["p"], wl["D"]))
    if adjoint:  # (NURBS: the weights are read too)
        return nsamp * nout * s + ncp * nout * s + tables + (ncp * s if wl["W"] is not None else 0)
    b = F * nsamp * nout * s + ncp * nout * s + tables
    if wl["W"] is not None:
        b += ncp * s
    return b


# --------------------------------------------------------------------------
# CPU baseline: the oracle's restatement of the reference algorithm
# --------------------------------------------------------------------------

def host_cpus():
    """(threads to use, description): every CPU this process may run on, bounded
    by the cgroup's CPU quota when one is set (a container's share of a larger
    host), plus nproc, os.cpu_count() and the CPU model for the record."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = max(1, min(affinity, quota) if quota else affinity)
    return threads, {"nproc": affinity, "os_cpu_count": os.cpu_count(),
                     "cgroup_cpu_quota": quota, "cpu_model": model}


def cpu_baseline(wl, budget_s: float = 16.0):
    """The reference algorithm (the oracle's restatement of the direct
    (p+1)^Nin sum per sample, src/spline_grid.jl:129-183) on the host: all
    usable cores in the workload's value type (the headline leg), one core,
    and all cores in FP64, each on slabs of the slowest sample axis for a
    share of ``budget_s``."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    nthreads, host = host_cpus()
    dims = []
    for i in range(len(wl["n"])):
        if wl["knots"] is not None:
            vals, mults = wl["knots"][i]
            knots = O.expand_knots(vals.astype(np.float64), mults)
        else:
            v, mu = O.default_knot_values(wl["n"][i], wl["p"][i], dtype=wl["dtype"])
            knots = O.expand_knots(v.astype(np.float64), mu)
        t = O.sample_points((knots[0], knots[-1]), wl["m"][i], wl["dtype"]).astype(np.float64)
        dims.append(O.Dim(knots, t, wl["p"][i], wl["D"][i]))
    row = math.prod(wl["m"][:-1])

    def leg(dtype, threads, seconds):
        P = wl["P"].astype(dtype)
        W = None if wl["W"] is None else wl["W"].astype(dtype)
        slab_rows = max(1, min(wl["m"][-1], 256 * threads * 4096 // max(1, row)))
        done, start, t0 = 0, 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            stop = min(start + slab_rows, wl["m"][-1])
            for f in wl["fields"]:
                O.eval_direct(dims, P, deriv=f, W=W, dtype=dtype,
                              last_range=(start, stop), nthreads=threads)
                done += row * (stop - start)
            start = 0 if stop >= wl["m"][-1] else stop
        dt = time.perf_counter() - t0
        return done / dt, done, dt, slab_rows

    v, done, dt, slab = leg(wl["dtype"], nthreads, budget_s / 2)
    v1, done1, dt1, _ = leg(wl["dtype"], 1, budget_s / 4)
    v64, done64, dt64, _ = leg(np.float64, nthreads, budget_s / 4)
    tname = np.dtype(wl["dtype"]).name
    return {"value": v, "unit": "sample-pts/s", "cores": nthreads, "kind": "port",
            "sample": f"the reference algorithm (direct (p+1)^Nin sum per sample, {tname}) "
                      f"restated in C, OpenMP over {nthreads} threads, on slabs of {slab} rows "
                      f"of the slowest sample axis of the same workload and inputs, all fields: "
                      f"{done} sample-points in {dt:.1f} s",
            "host": host,
            "one_core": {"value": v1, "dtype": tname,
                         "sample": f"{done1} sample-points in {dt1:.1f} s on 1 thread"},
            "fp64_all_cores": {"value": v64, "dtype": "float64",
                               "sample": f"{done64} sample-points in {dt64:.1f} s on "
                                         f"{nthreads} threads"}}


def fp64_errors(sg, grid, wl, outs):
    """Max and RMS |result - FP64 result| of each field over every sample,
    relative to max|FP64 field|.  The FP64 result is this library's FP64 path
    on the same knots, sample points, and FP32-rounded control points and
    weights promoted exactly to FP64; the GPU tests hold that path to 1e-12 of
    the CPU oracle.  Returns (max per field, RMS per field)."""
    import torch

    g64 = fp64_grid(sg, grid)
    errs, rms = [], []
    ref = sg.colmajor_empty(g64.eval.shape, torch.float64, "cuda")
    for f, out in zip(wl["fields"], outs):
        if wl.get("rational"):
            sg.evaluate_rational_fields_(g64, [f], [ref])
        else:
            sg.evaluate_(g64, derivative_order=f, eval=ref)
        d = out.double() - ref
        scale = float(ref.abs().max()) or 1.0  # an all-zero field: absolute error
        errs.append(float(d.abs().max()) / scale)
        rms.append(float(d.pow(2).mean().sqrt()) / scale)
        del d
    return errs, rms


def batch_fp64_errors(grids, refs):
    """([max over grids of max|eval - FP64| / max|FP64|], [RMS over all samples
    of |eval - FP64| / max|FP64| of the sample's grid])."""
    worst, sq, count = 0.0, 0.0, 0
    for g, ref in zip(grids, refs):
        scale = float(ref.abs().max()) or 1.0
        d = (g.eval.double() - ref).abs() / scale
        worst = max(worst, float(d.max()))
        sq += float(d.pow(2).sum())
        count += d.numel()
    return [worst], [math.sqrt(sq / max(count, 1))]


def fp64_grid(sg, grid):
    """This library's FP64 grid on the same knots, sample points and
    FP32-rounded control points (and weights) promoted exactly to FP64."""
    dims64 = tuple(sg.SplineDimension(
        s.n_basis_functions, s.degree, s.n_sample_points,
        max_derivative_order=s.max_derivative_order, float_type=np.float64,
        knot_vector=sg.KnotVector(s.knot_vector.knot_values.astype(np.float64),
                                  s.knot_vector.multiplicities),
        sample_points=s.host_sample_points(), device="cuda") for s in grid.spline_dimensions)
    g64 = (sg.NURBSGrid if grid.is_nurbs else sg.SplineGrid)(dims64, grid.Nout)
    sg.obtain(g64.control_points).copy_(sg.obtain(grid.control_points).double())
    if grid.is_nurbs:
        g64.weights.copy_(grid.weights.double())
    return g64


def fp64_adjoint_errors(sg, grid, ebar, pbar):
    """Max and RMS |Pbar - FP64 Pbar| relative to max|FP64 Pbar|, the FP64
    adjoint taking the same ebar promoted exactly; returned as ([max], [RMS])
    like fp64_errors' per-field lists."""
    import torch

    g64 = fp64_grid(sg, grid)
    ref = sg.colmajor_zeros(tuple(pbar.shape), torch.float64, "cuda")
    if g64.is_nurbs:
        sg.evaluate_rational_adjoint_(g64, eval=sg.colmajor(ebar.double()), control_points=ref)
    else:
        sg.evaluate_adjoint_(g64, eval=sg.colmajor(ebar.double()), control_points=ref)
    d = pbar.double() - ref
    scale = float(ref.abs().max()) or 1.0
    return [float(d.abs().max()) / scale], [float(d.pow(2).mean().sqrt()) / scale]


def traffic_from_profiles(workload_name: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{workload_name}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes_per_launch")


# --------------------------------------------------------------------------
# main
# --------------------------------------------------------------------------

def store_ceiling():
    """Write bandwidth of the 805 MB of the 512 x 512 x 256 x 3 forward as a
    plain fill and in the forward kernel's store pattern with nothing
    computed (splinegrids.jl_amd/csrc/store_ceiling.hip, built by build()):
    the ceilings its roofline fraction is read against."""
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "splinegrids.jl_amd",
                       "store_ceiling")
    if not os.path.exists(exe):
        print("store_ceiling is not built: run __graft_entry__.build()", file=sys.stderr)
        return 1
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    return r.returncode


def copy_calibration(args, wl):
    """Time torch's device copy of 1 GiB; its counters calibrate FETCH/WRITE_SIZE."""
    import torch

    src = torch.ones(wl["m"][0], dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    fill = wl["copy"] == "fill"
    op = (lambda: dst.fill_(1.5)) if fill else (lambda: dst.copy_(src))
    for _ in range(args.warmup):
        op()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        op()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    nbytes = (1 if fill else 2) * src.numel() * 4
    print(json.dumps({"metric": f"{wl['copy']} GB/s", "value": nbytes / (ms * 1e-3) / 1e9,
                      "unit": "GB/s", "ms_per_step": ms, "bytes_per_step": nbytes,
                      "config": {"workload": wl["name"]}}), flush=True)
    return 0


def batch_bytes(grids) -> int:
    total = 0
    for g in grids:
        nsamp = math.prod(s.n_sample_points for s in g.spline_dimensions)
        ncp = math.prod(s.n_basis_functions for s in g.spline_dimensions)
        total += 4 * (nsamp * g.Nout + ncp * g.Nout + (ncp if g.is_nurbs else 0))
        total += sum(s.n_sample_points * (s.degree + 1) * 4 + 4 * s.n_sample_points
                     for s in g.spline_dimensions)
    return total


def settle_gc() -> None:
    """Collect once and freeze what set-up left (torch, the grids), so that no
    full collection scans those objects in the middle of the timed loop; the
    collector stays on for the loop's own objects.  A full collection after
    building a large grid takes tens of milliseconds, more than a hundred
    launches of a short kernel."""
    gc.collect()
    gc.freeze()


def timed_steps(step, steps: int, warmup: int):
    """Average ms per step over `steps` back-to-back steps on the current
    stream (HIP events), after `warmup` untimed steps; also the wall time.
    The collector is settled before the warm-up, so that the GPU goes from
    the warm-up straight into the timed steps."""
    import torch

    settle_gc()
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        step()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps, time.perf_counter() - t_wall


def bench_batch(args, sg, wl):
    """The 1024-grid batch: the per-grid launch loop, the batched persistent launch, and the
    batched launch captured in a hipGraph; value is the graph replay."""
    import torch

    rng = np.random.default_rng(SEED0 + 5)
    specs = c5_specs(rng)
    t0 = time.perf_counter()
    grids = [build_spec_grid(sg, spec, rng) for spec in specs]
    batch = sg.SplineGridBatch(grids)
    setup_s = time.perf_counter() - t0
    info = batch.autotune(reps=3)

    def timed(fn, steps):
        settle_gc()
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    steps = max(5, args.steps // 10)
    ms_loop = timed(batch.evaluate_each_, steps)
    ms_batch = timed(batch.evaluate_, steps)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        batch.evaluate_()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=stream):
            batch.evaluate_()
    torch.cuda.synchronize()
    ms_graph = timed(graph.replay, steps)
    nsamp = info["samples"]
    B = batch_bytes(grids)
    # errors of every grid against this library's FP64 path on the same inputs,
    # for the native batch (the graph replays' result) and the MFMA tiers
    refs = []
    for g in grids:
        g64 = fp64_grid(sg, g)
        sg.evaluate_(g64)
        refs.append(g64.eval)
    errs = batch_fp64_errors(grids, refs)
    tiers = {"native": {"ms_per_step": ms_graph, "value": nsamp / (ms_graph * 1e-3),
                        "max_rel_err_vs_fp64": errs[0], "rms_rel_err_vs_fp64": errs[1]}}
    if not args.no_tiers:
        for tier in ("split_f16", "bf16"):
            try:
                t_ms = timed(lambda: batch.evaluate_(tier=tier), steps)
            except sg.SplineGridsError as e:
                tiers[tier] = {"unsupported": str(e)}
                continue
            torch.cuda.synchronize()
            e_max, e_rms = batch_fp64_errors(grids, refs)
            tiers[tier] = {"ms_per_step": t_ms, "value": nsamp / (t_ms * 1e-3),
                           "max_rel_err_vs_fp64": e_max, "rms_rel_err_vs_fp64": e_rms}
    out = {
        "metric": METRIC, "value": nsamp / (ms_graph * 1e-3), "unit": "sample-pts/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms_graph,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded inputs of the workload's shape)",
        "config": {"workload": wl["name"], "grids": len(grids), **info,
                   "describe": batch.describe(),
                   "setup_s": setup_s},
        "roofline": {"bound": "hbm", "achieved": B / (ms_graph * 1e-3) / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": B / (ms_graph * 1e-3) / 1e9 / PEAK_HBM_GBS, "traffic": traffic_from_profiles(args.workload),
                     "algorithmic_bytes_per_step": B},
        "per_grid_loop_ms": ms_loop, "batched_ms": ms_batch, "graph_ms": ms_graph,
        "speedup_vs_per_grid_loop": ms_loop / ms_graph,
        "max_rel_err_vs_fp64": errs[0], "rms_rel_err_vs_fp64": errs[1], "tiers": tiers,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--tier", default="native")
    ap.add_argument("--cpu-budget", type=float, default=16.0,
                    help="seconds of CPU-baseline work over its three legs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tiers", action="store_true",
                    help="skip timing the split-FP16 and BF16 tiers beside FP32")
    ap.add_argument("--separate-fields", action="store_true",
                    help="one launch per field instead of the fused launch")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the evaluation shards by sample slabs with no "
              "communication, so a multi-GPU figure is N independent single-GPU runs; "
              "use --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import __graft_entry__

    if args.workload == "ceiling":
        return store_ceiling()
    sg = __graft_entry__.package()
    wl = workload(args.workload)
    if wl.get("copy"):
        return copy_calibration(args, wl)
    if wl.get("batch"):
        return bench_batch(args, sg, wl)
    grid = build_grid(sg, wl)
    adjoint = wl.get("adjoint", False)
    fields = wl["fields"]
    outs = [grid.eval] + [torch.empty_like(grid.eval) for _ in fields[1:]]
    if adjoint:
        ebar = sg.colmajor(torch.randn(grid.eval.shape, dtype=grid.dtype, device="cuda"))
        pbar = sg.colmajor_zeros(grid.control_points.shape, grid.dtype, "cuda")

    def make_step(tier):
        def step():
            if wl.get("rational"):
                sg.evaluate_rational_fields_(grid, fields, outs)
            elif adjoint and grid.is_nurbs:
                sg.evaluate_rational_adjoint_(grid, eval=ebar, control_points=pbar)
            elif adjoint:
                sg.evaluate_adjoint_(grid, eval=ebar, control_points=pbar, tier=tier)
            elif len(fields) == 1:
                sg.evaluate_(grid, derivative_order=fields[0], tier=tier)
            elif args.separate_fields:
                for f, o in zip(fields, outs):
                    sg.evaluate_(grid, derivative_order=f, eval=o, tier=tier)
            else:
                sg.evaluate_fields_(grid, fields, outs, tier=tier)
        return step

    ms, t_wall = timed_steps(make_step(args.tier), args.steps, args.warmup)
    nsamp = math.prod(wl["m"])
    n_fields = 1 if adjoint else len(fields)
    value = n_fields * nsamp / (ms * 1e-3)
    B = algorithmic_bytes(wl, adjoint)
    achieved = B / (ms * 1e-3) / 1e9
    launches = 1 if (adjoint or len(fields) == 1 or not args.separate_fields) else len(fields)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "sample-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if wl["dtype"] == np.float64 else "f32",
        "data": "synthetic (seeded inputs of the workload's shape)",
        "config": {"workload": wl["name"], "tier": args.tier, "fields": len(fields),
                   "fused_fields": not args.separate_fields, "launches_per_step": launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic_from_profiles(args.workload),
                     "algorithmic_bytes_per_step": B},
        "wall_s_timed": t_wall,
    }
    if not adjoint:
        if wl["dtype"] != np.float64:
            # an FP64 workload would be compared with itself: no figure for it
            out["max_rel_err_vs_fp64"], out["rms_rel_err_vs_fp64"] = fp64_errors(sg, grid, wl, outs)
    else:
        out["max_rel_err_vs_fp64"], out["rms_rel_err_vs_fp64"] = fp64_adjoint_errors(
            sg, grid, ebar, pbar)
    # the other arithmetic tiers on the same inputs, timed the same way
    if (wl["dtype"] == np.float32 and args.tier == "native" and not args.no_tiers
            and not wl.get("rational") and not (adjoint and wl["W"] is not None)):
        tiers = {"native": {"ms_per_step": ms, "value": value,
                            "max_rel_err_vs_fp64": out["max_rel_err_vs_fp64"],
                            "rms_rel_err_vs_fp64": out["rms_rel_err_vs_fp64"]}}
        for tier in ("split_f16", "bf16"):
            try:
                t_ms, _ = timed_steps(make_step(tier), max(20, args.steps // 2), args.warmup)
            except sg.SplineGridsError as e:
                tiers[tier] = {"unsupported": str(e)}
                continue
            mx, rms = (fp64_adjoint_errors(sg, grid, ebar, pbar) if adjoint
                       else fp64_errors(sg, grid, wl, outs))
            tiers[tier] = {"ms_per_step": t_ms, "value": n_fields * nsamp / (t_ms * 1e-3),
                           "max_rel_err_vs_fp64": mx, "rms_rel_err_vs_fp64": rms}
        out["tiers"] = tiers
    if adjoint:
        # forward (evaluate!) plus backward (evaluate_adjoint!) per step
        def fwd_bwd():
            sg.evaluate_(grid)
            if grid.is_nurbs:
                sg.evaluate_rational_adjoint_(grid, eval=ebar, control_points=pbar)
            else:
                sg.evaluate_adjoint_(grid, eval=ebar, control_points=pbar)
        out["forward_plus_adjoint_ms"], _ = timed_steps(fwd_bwd, max(20, args.steps // 2),
                                                        args.warmup)
    if not args.no_cpu_baseline and not wl.get("rational") and not (adjoint and wl["W"] is not None):
        # (the reference has no NURBS derivatives to time)
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_budget)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
