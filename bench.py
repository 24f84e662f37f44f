"""Benchmark: 3D 7-point JACOBI sweep, 512^3 FP32, radius 1, 100 sweeps.

One "step" is one pass of the workload: 100 sweeps of the 512^3 grid, with
the field already resident in HBM (seeded boundary shell written on the GPU
once, zero interior; each step continues from the previous state).  Prints
one JSON line.

  value      GCells/s = 510^3 interior cells x 100 sweeps x steps / time
  roofline   dominant kernel vs the 8 TB/s HBM peak, algorithmic bytes
             4 B x (512^3 + 510^3) per sweep, launch duration from HIP events
             on the launch stream; frac_per_hbm_pass divides the bytes one
             HBM pass must move (the same figure, once per launch of two
             fused sweeps) by the launch duration
  cpu_baseline  the CPU oracle (C, OpenMP) on a bounded sample of the same
             workload, on this host's cores
  legs       the other measured workloads, each with its own roofline:
             mixed384 (384^3 R=4, 200 sweeps in FP32, FP64, BF16 and FP16
             storage with defect correction; errors against the FP64 CPU
             oracle, whose run is that leg's CPU baseline) and batched (512
             uneven grids, 50 sweeps, one graph launch per sweep against
             per-problem launches) and reference_sizes (the reference's own
             evaluation sizes 1024^3 and 2D 32768^2); --no-legs skips them
  warm_start the metric's second half: time to relative residual 1e-6 at
             128^3, learned warm start over plain zero start, on two sets of
             five problems (--no-warm skips it)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 512
RADIUS = 1
SWEEPS = 100
SEED = 0x5EED0002
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_sweep(n=N, r=RADIUS, es=4):
    return es * (n ** 3 + (n - 2 * r) ** 3)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def warm_start_leg(problems: int = 5, coarse_sweeps: int = 2000):
    """The metric's second half: wall time to relative residual 1e-6 on
    128^3 Laplace problems with Fourier-mode faces, learned warm start
    against the plain zero start.  The pipeline of record is
    "learned_coarse": the tuned U-Net's 64^3 prediction refined by 2000
    JACOBI sweeps at 64^3, then upsampled; "learned" (the prediction
    upsampled directly) and "blend_coarse" (the same refinement of the
    network's input guess, the control without the network) are reported
    beside it.  Two sets of `problems` each: the five the network's training
    was judged on, and five further ones never used in any choice."""
    import numpy as np
    from thesis_amd import warmstart

    weights = os.path.join(ROOT, "thesis_amd", "tuned", "unet64.npz")
    with np.load(weights, allow_pickle=False) as z:
        state = {k: z[k] for k in z.files}
    rows = []
    for label, base in (("standard", 0x5EED0005), ("further", 0x5EED0100)):
        for i in range(problems):
            r = warmstart.compare_warm_start(state, seed=base + 1_000_000 + i, n=128, m=64, tol=1e-6,
                                             check_every=200, coarse_sweeps=coarse_sweeps)
            rows.append({"set": label, "plain_sweeps": r["plain"]["sweeps"],
                         "plain_s": round(r["plain"]["seconds"], 4), "learned_sweeps": r["learned"]["sweeps"],
                         "learned_s": round(r["learned"]["seconds"], 4),
                         "network_s": round(r["learned"]["network_seconds"], 5),
                         "learned_coarse_sweeps": r["learned_coarse"]["sweeps"],
                         "learned_ratio": round(r["time_ratio_learned_vs_plain"], 4),
                         "learned_coarse_ratio": round(r["time_ratio_learned_coarse_vs_plain"], 4),
                         "blend_coarse_ratio": round(r["time_ratio_blend_coarse_vs_plain"], 4)})
    mean = lambda k, sel=None: round(float(np.mean([r[k] for r in rows if sel is None or r["set"] == sel])), 4)
    return {"metric": "time to relative residual 1e-6 at 128^3: learned warm start / plain zero start",
            "unit": "ratio", "higher_is_better": False, "problems": len(rows),
            "of_record": "learned_coarse",
            "value": mean("learned_coarse_ratio"),
            "value_max": round(max(r["learned_coarse_ratio"] for r in rows), 4),
            "learned_coarse": mean("learned_coarse_ratio"), "blend_coarse": mean("blend_coarse_ratio"),
            "learned": mean("learned_ratio"), "learned_standard": mean("learned_ratio", "standard"),
            "learned_further": mean("learned_ratio", "further"),
            "learned_max": round(max(r["learned_ratio"] for r in rows), 4),
            "network_seconds": mean("network_s"),
            "coarse_sweeps": coarse_sweeps, "per_problem": rows}


def cpu_baseline(budget_s: float = 12.0, single_budget_s: float = 6.0):
    """Time the CPU oracle on sweeps of the same field (FP32) until the
    budget is used (at least 2 sweeps): OpenMP over all host threads (the
    reported value), then one thread (the reference's own CPU loop,
    src/stencil_cpu.cu:32-35, is single-threaded); GCells/s."""
    from oracle import oracle as O

    def timed(threads, budget):
        u1 = O.init_seeded(N, N, N, 3, RADIUS, SEED)
        u2 = u1.copy()
        w = O.default_weight(3, RADIUS)
        sweeps = 0
        t0 = time.perf_counter()
        while True:
            O.sweep(u1, u2, 3, RADIUS, O.OP_JACOBI, w, threads)
            u1, u2 = u2, u1
            sweeps += 1
            el = time.perf_counter() - t0
            if (el >= budget and sweeps >= 2) or sweeps >= SWEEPS:
                break
        return sweeps, el, (N - 2 * RADIUS) ** 3 * sweeps / el / 1e9

    threads = O.num_threads()
    sweeps, el, rate = timed(threads, budget_s)
    sweeps1, el1, rate1 = timed(1, single_budget_s)
    return {"value": round(rate, 4), "unit": "GCells/s", "cores": threads,
            "kind": "port",
            "sample": f"{sweeps} sweeps of the {N}^3 R={RADIUS} FP32 JACOBI field "
                      f"({el:.1f} s, oracle/stencil_oracle.c, OpenMP {threads} threads)",
            "single_thread": {"value": round(rate1, 4), "cores": 1,
                              "sample": f"{sweeps1} sweeps, {el1:.1f} s"},
            "cpu_model": _cpu_model()}


def load_traffic(tag: str):
    """HBM bytes per launch of the dominant kernel from a committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (calibrated against the
    copy kernel), with the file it came from."""
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        return d.get("bytes_per_launch"), f"profiles/traffic_{tag}.json ({d.get('kernel', '?')}: " \
                                          f"{d.get('workload', '?')})"
    return None, None


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2] if len(xs) % 2 else 0.5 * (xs[len(xs) // 2 - 1] + xs[len(xs) // 2])


def batched_leg(st, np, repeats: int = 3):
    """512 uneven 3D grids (edges log-uniform in [16, 128], radius 1 or 2),
    50 JACOBI sweeps each: one hipGraph of one launch per pass over all grids
    (two fused sweeps per pass where the shells are equal), one persistent
    cooperative launch of the same passes captured in a hipGraph, and
    per-problem launches.  The roofline is the graph
    mode's: algorithmic bytes of one sweep of every grid (4 B x (cells +
    interior cells)) over the time per sweep."""
    from thesis_amd import workloads
    from thesis_amd.tuner.batched_tiles import load_rules

    shapes = workloads.batched_shapes(512)
    its = 50
    rules = load_rules()
    stream = st.Stream()
    probs, bufs = [], []
    for i, (nx, ny, nz, r) in enumerate(shapes):
        p = st.StencilParams.make(nx=nx, ny=ny, nz=nz, dimensions=3, radius=r, op=st.OP_JACOBI)
        d1 = st.DeviceArray(p.shape, np.float32)
        d2 = st.DeviceArray(p.shape, np.float32)
        st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=0x5EED0004 + i, stream=stream)
        probs.append(dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=its, op=st.OP_JACOBI, u1=d1, u2=d2))
        bufs += [d1, d2]
    stream.sync()
    interior = workloads.batched_interior_cells(shapes)
    bytes_sweep = 4 * (workloads.batched_cells(shapes) + interior)
    res = {}
    for name, mode in (("graph", st.BATCH_GRAPH), ("persistent", st.BATCH_PERSISTENT),
                       ("sequential", st.BATCH_SEQUENTIAL)):
        # the first dispatch in graph mode captures and instantiates the
        # graph (counted in its time); later ones replay it
        first = st.dispatch_batched(probs, rules, mode, stream=stream, timed=True)
        ts = [st.dispatch_batched(probs, rules, mode, stream=stream, timed=True) for _ in range(repeats)]
        t = _median(ts)
        res[name] = {"ms": round(t, 3), "gcells_per_s": round(interior * its / t / 1e6, 2),
                     "hbm_frac": round(bytes_sweep * its / (t / 1e3) / 8e12, 4),
                     "first_call_ms": round(first, 3)}
    for b in bufs:
        b.free()
    per_sweep_s = res["graph"]["ms"] / 1e3 / its
    achieved = bytes_sweep / per_sweep_s / 1e9
    return {"metric": "batched GCells/s (512 uneven grids, 50 sweeps)", "value": res["graph"]["gcells_per_s"],
            "unit": "GCells/s", "higher_is_better": True, "dtype": "f32",
            "data": "synthetic: 512 grids, edges round(16*8^U) per axis, radius 1|2, seeded boundaries",
            "config": {"workload": "batched uneven 3D JACOBI, 50 sweeps", "total_cells": workloads.batched_cells(shapes),
                       "tile_rules": rules},
            "modes": res,
            "speedup_graph_vs_sequential": round(res["sequential"]["ms"] / res["graph"]["ms"], 2),
            "speedup_persistent_vs_sequential": round(res["sequential"]["ms"] / res["persistent"]["ms"], 2),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "bytes_per_sweep_algorithmic": bytes_sweep,
                         "ms_per_sweep": round(per_sweep_s * 1e3, 5),
                         "note": "per-sweep convention over the whole graph (graph time / 50 sweeps): the "
                                 "graph runs 24 passes of two fused sweeps (equal shells) and 2 single-sweep "
                                 "passes, one launch each, so `frac` counts two sweeps per fused HBM pass; "
                                 "the time includes the shell comparison and its read-back, once per "
                                 "dispatch; modes.graph.first_call_ms adds capture and instantiation"}}


def mixed_leg(st, np, repeats: int = 3, with_cpu: bool = True):
    """384^3, radius 4 (25-point), JACOBI c_r = 1/24, 200 sweeps: FP32, FP64,
    and BF16 / FP16 storage with FP32 arithmetic and FP32 defect correction
    every 10 sweeps.  Errors (max-abs, relative L2) are taken against the
    FP64 CPU oracle run of the same 200 sweeps, which is also this leg's CPU
    baseline.  The roofline is the BF16 correction sweep's, the leg's
    dominant kernel: per-sweep algorithmic bytes 2 B x (cells + interior
    cells), plus the defect read at every interior cell in the
    kernel-bytes figure; its time per sweep is (t(200) - t(100)) / 100 of
    dispatches without defect correction, which cancels the fixed passes."""
    n, r, its = 384, 4, 200
    seed = 0x5EED0003
    base = dict(dim=n, radius=r, iterations=its, op=st.OP_JACOBI, heuristic=1)
    shape = (n, n, n)
    cells, interior = n ** 3, (n - 2 * r) ** 3
    out = {"metric": "GCells/s and error vs the FP64 CPU oracle per storage type (384^3 r=4, 200 sweeps)",
           "unit": "GCells/s", "higher_is_better": True,
           "config": {"workload": "3D 25-point JACOBI u'=(1/24)*sum of 24 neighbours, 384^3, radius 4, "
                                  "200 sweeps; 16-bit storage with FP32 defect correction every 10 sweeps"},
           "data": "synthetic: seeded uniform [0,1) boundary shell (4 cells thick), zero interior",
           "results": {}}
    ref = None
    if with_cpu:
        from oracle import oracle as O
        x64 = O.init_seeded(n, n, n, 3, r, seed, dtype=np.float64)
        t0 = time.perf_counter()
        ref = O.run(x64, x64.copy(), 3, r, O.OP_JACOBI, its).copy()
        el = time.perf_counter() - t0
        threads = O.num_threads()
        out["cpu_baseline"] = {"value": round(interior * its / el / 1e9, 4), "unit": "GCells/s",
                               "cores": threads, "kind": "port",
                               "sample": f"the whole leg in FP64: {its} sweeps of the {n}^3 R={r} JACOBI field "
                                         f"({el:.1f} s, oracle/stencil_oracle.c, OpenMP {threads} threads); "
                                         f"its field is the error reference"}
        del x64

    def rec(ms, es_bytes, u=None):
        d = {"ms": round(ms, 3), "gcells_per_s": round(interior * its / ms / 1e6, 1),
             "hbm_frac": round(es_bytes * (cells + interior) * its / (ms / 1e3) / 8e12, 4)}
        if u is not None and ref is not None:
            diff = u.astype(np.float64) - ref
            d["max_abs"] = float(np.max(np.abs(diff)))
            d["l2_rel"] = float(np.sqrt(np.sum(diff * diff)) / np.sqrt(np.sum(ref * ref)))
            d["bitwise_equal_to_oracle"] = bool(u.dtype == np.float64 and np.array_equal(
                u.view(np.uint64), ref.view(np.uint64)))
        return d

    for name, dt, npdt, es in (("f64", st.F64, np.float64, 8), ("f32", st.F32, np.float32, 4)):
        p = st.StencilParams.make(dtype=dt, variant="auto", **base)
        a = st.DeviceArray(shape, npdt)
        b = st.DeviceArray(shape, npdt)
        ts = []
        for _ in range(repeats + 1):
            st.initialize_region(p, a, b, st.INIT_SEEDED, seed=seed)
            _, ms = st.dispatch_kernels(p, a, b, timed=True)
            ts.append(ms)
        out["results"][name] = rec(_median(ts[1:]), es, a.numpy())
        if name == "f32":  # the initial field of the 16-bit runs' FP32 base
            st.initialize_region(p, a, b, st.INIT_SEEDED, seed=seed)
            x0 = a.numpy()
        a.free()
        b.free()
    bufs = {}
    for name, dt in (("bf16", st.BF16), ("f16", st.F16)):
        pm = st.StencilParams.make(dtype=dt, restart=10, **base)
        x = st.DeviceArray(shape, np.float32)
        e1, e2, d = (st.DeviceArray(shape, np.uint16) for _ in range(3))
        u = st.DeviceArray(shape, np.float32)
        ts = []
        for _ in range(repeats + 1):
            x.upload(x0)
            ts.append(st.dispatch_mixed(pm, x, e1, e2, d, u, timed=True))
        out["results"][name] = rec(_median(ts[1:]), 2, u.numpy())
        out["results"][name]["restart"] = 10
        bufs[name] = (pm, x, e1, e2, d, u)
    # the BF16 correction sweep alone: dispatches with no defect correction
    pm, x, e1, e2, d, u = bufs["bf16"]
    per = []
    for _ in range(repeats):
        tt = {}
        for k in (200, 100):
            q = pm.copy(iterations=k, restart=k)
            x.upload(x0)
            tt[k] = st.dispatch_mixed(q, x, e1, e2, d, u, timed=True)
        per.append((tt[200] - tt[100]) / 100.0)
    sweep_ms = _median(per)
    alg = 2 * (cells + interior)
    kern = alg + 2 * interior
    achieved = alg / (sweep_ms / 1e3) / 1e9
    out["roofline"] = {"bound": "hbm", "kernel": "BF16 correction sweep (row kernel, 16-bit mode)",
                       "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "bytes_per_launch_algorithmic": alg, "bytes_per_launch_kernel": kern,
                       "frac_kernel_bytes": round(kern / (sweep_ms / 1e3) / 8e12, 4),
                       "avg_launch_ms": round(sweep_ms, 5),
                       "note": "algorithmic bytes 2 B x (cells + interior) per sweep; the kernel also reads "
                               "the 16-bit defect at every interior cell (bytes_per_launch_kernel)"}
    for t in bufs.values():
        for b in t[1:]:
            b.free()
    out["value"] = out["results"]["bf16"]["gcells_per_s"]
    return out


def size_legs(st, np, repeats: int = 3):
    """The reference's own evaluation sizes and sweep count
    (scripts/evaluate_stencils.py:9,71-76): 3D 1024^3 and 2D 32768^2, 8
    sweeps, R=1 FP32 JACOBI, seeded boundary, AUTO's choice of kernel;
    GCells/s and the per-sweep algorithmic HBM fraction (4 B x (cells +
    interior cells) per sweep)."""
    out = {}
    for name, dims, n in (("sweep1024", 3, 1024), ("sweep2d_32768", 2, 32768)):
        its = 8
        p = st.StencilParams.make(nx=n, ny=n, nz=n if dims == 3 else 1, dimensions=dims, radius=1,
                                  iterations=its, op=st.OP_JACOBI, variant="auto", heuristic=1)
        cfg = st.launch_config(p)
        a = st.DeviceArray(p.shape, np.float32)
        b = st.DeviceArray(p.shape, np.float32)
        st.initialize_region(p, a, b, st.INIT_SEEDED, seed=SEED)
        st.dispatch_kernels(p, a, b)  # warm-up
        ts = []
        for _ in range(repeats):
            _, ms = st.dispatch_kernels(p, a, b, timed=True)
            ts.append(ms)
        a.free()
        b.free()
        ms = _median(ts)
        cells = n ** dims
        interior = (n - 2) ** dims
        out[name] = {"workload": f"{dims}D {n}^{dims} R=1 FP32 JACOBI, {its} sweeps",
                     "ms": round(ms, 3), "gcells_per_s": round(interior * its / ms / 1e6, 1),
                     "hbm_frac": round(4 * (cells + interior) * its / (ms / 1e3) / 8e12, 4),
                     "variant": {v: k for k, v in st.VARIANTS.items()}[cfg["variant"]], "launch": cfg}
    return out


def probe(st, np):
    d1 = st.DeviceArray((N, N, N), np.float32)
    d2 = st.DeviceArray((N, N, N), np.float32)
    p = st.StencilParams.make(dim=N, radius=RADIUS, iterations=2, op=st.OP_JACOBI, heuristic=1)
    st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=SEED)
    st.copy_f32(d2, d1, N ** 3)
    st.synchronize()
    print(f"[probe] copy {N}^3 fp32 ({4 * N ** 3} bytes read, {4 * N ** 3} written)", flush=True)
    for v, its, form, per_pass in (("smem_register", 2, 0, 0), ("base", 2, 0, 0), ("smem_padded", 2, 0, 0),
                                   ("coop", 2, 1, 0), ("coop", SWEEPS, 3, 1)):
        q = p.copy(variant=v, iterations=its, temporal_form=form, launch_per_pass=per_pass)
        st.initialize_region(q, d1, d2, st.INIT_SEEDED, seed=SEED)
        st.dispatch_kernels(q, d1, d2)
        st.synchronize()
        print(f"[probe] {v} form {form}: {its} sweeps, launch {st.launch_config(q)}", flush=True)


def headline_leg(st, np, args):
    """The 512^3 sweep (or the grid the --dim/--radius/--sweeps/--dtype
    knobs name): K timed steps of `sweeps` sweeps each after W warm-up
    steps, synchronised on both sides."""
    headline = (args.dim, args.radius, args.sweeps, args.dtype) == (N, RADIUS, SWEEPS, "f32")
    n, radius, sweeps = args.dim, args.radius, args.sweeps
    if sweeps % 2:
        raise SystemExit("--sweeps must be even (the state stays in the first buffer)")
    f64 = args.dtype == "f64"
    es = 8 if f64 else 4
    kw = dict(dim=n, radius=radius, iterations=sweeps, op=st.OP_JACOBI, dtype=st.F64 if f64 else st.F32,
              variant=args.variant, heuristic=1, temporal_steps=args.temporal_steps)
    if args.block:
        bx, by, bz = (int(v) for v in args.block.split(","))
        kw.update(block_x=bx, block_y=by, block_z=bz, heuristic=0)
    if args.coarsen:
        cx, cy = (int(v) for v in args.coarsen.split(","))
        kw.update(coarsen_x=cx, coarsen_y=cy)
    p = st.StencilParams.make(**kw)
    cfg = st.launch_config(p)
    stream = st.Stream()
    d1 = st.DeviceArray(p.shape, np.float64 if f64 else np.float32)
    d2 = st.DeviceArray(p.shape, np.float64 if f64 else np.float32)

    # The field is initialised once; every step advances it by `sweeps` more
    # sweeps (an even count, so the state stays in d1).  The work per sweep
    # does not depend on the values.
    st.initialize_region(p, d1, d2, st.INIT_SEEDED if args.init == "seeded" else st.INIT_REFERENCE,
                         seed=SEED, stream=stream)

    def step():
        in_u2, ms = st.dispatch_kernels(p, d1, d2, stream=stream, timed=True)
        assert not in_u2
        return ms

    for _ in range(args.warmup):
        step()
    stream.sync()
    st.synchronize()
    t0 = time.perf_counter()
    sweep_ms = [step() for _ in range(args.steps)]
    stream.sync()
    st.synchronize()
    wall = time.perf_counter() - t0

    sweep_total_s = sum(sweep_ms) / 1e3
    cells = (n - 2 * radius) ** 3 * sweeps * args.steps
    gcells = cells / wall / 1e9
    launches = cfg["launches"]
    bytes_sweep = algorithmic_bytes_per_sweep(n, radius, es)
    bytes_launch = bytes_sweep * sweeps / launches
    avg_launch_s = sweep_total_s / (launches * args.steps)
    achieved = bytes_launch / avg_launch_s / 1e9
    variant_name = {v: k for k, v in st.VARIANTS.items()}[cfg["variant"]]
    # the committed traffic summaries were measured on the seeded 512^3 field
    traffic, traffic_src = load_traffic(variant_name) if headline and args.init == "seeded" else (None, None)
    fused = variant_name == "coop" and cfg["temporal_steps"] > 1
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch_algorithmic": bytes_launch,
            "avg_launch_ms": round(avg_launch_s * 1e3, 5)}
    if traffic_src:
        roof["traffic_source"] = traffic_src
    if fused:
        # one launch is one HBM pass that advances temporal_steps sweeps: the
        # bytes a pass must move are one sweep's (read the grid, write the
        # interior), so this is the fraction of the HBM peak the pass uses
        roof["bytes_per_hbm_pass"] = bytes_sweep
        roof["frac_per_hbm_pass"] = round(bytes_sweep / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4)
        roof["note"] = (f"a launch is one HBM pass of {cfg['temporal_steps']} fused sweeps, so `frac` (the "
                        f"per-sweep algorithmic convention) can exceed 1; frac_per_hbm_pass is the physical "
                        f"fraction; the event time per launch includes the shell comparison and the band "
                        f"fix-up launches (which return at once with equal shells)")
    else:
        roof["note"] = "one sweep per launch"
    out = {
        "metric": "GCells/s & %HBM roofline (512^3 r=1)" if headline
                  else f"GCells/s & %HBM roofline ({n}^3 r={radius} {args.dtype})",
        "value": round(gcells, 3),
        "unit": "GCells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "event_ms_per_step": round(sum(sweep_ms) / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": ("synthetic: seeded uniform [0,1) boundary shell, zero interior" if args.init == "seeded"
                 else "synthetic: reference initial condition (u1 shell 1, u2 zero)"),
        "config": {"workload": (f"3D {6 * radius + 1}-point JACOBI u'=(1/{6 * radius})*sum of "
                                f"{6 * radius} neighbours, {n}^3 {args.dtype.upper()}, radius {radius}, "
                                f"{sweeps} sweeps per step"),
                   "variant": variant_name, "launch": cfg, "device": st.device_name(0)},
        "roofline": roof,
    }
    d1.free()
    d2.free()
    stream.close()
    return out, headline


def main():
    global N, RADIUS, SWEEPS
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--variant", default="auto")
    ap.add_argument("--temporal-steps", type=int, default=0)
    ap.add_argument("--block", default="", help="bx,by,bz override (heuristic otherwise)")
    ap.add_argument("--coarsen", default="", help="cx,cy override")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-warm", action="store_true",
                    help="skip the learned warm-start leg of the headline line")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the mixed384 and batched legs of the headline line")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--workload", default="sweep512",
                    choices=["sweep512", "batched", "mixed384", "sizes"],
                    help="sweep512: the headline 512^3 sweep with its legs (default); batched / "
                         "mixed384 / sizes (1024^3 and 32768^2, 8 sweeps): that leg alone, as its "
                         "own line")
    ap.add_argument("--probe", action="store_true",
                    help="short run for counter collection: one streaming copy of a 512^3 "
                         "field, 2 sweeps of each single-sweep family and of the LDS-ring "
                         "COOP form, then 100 sweeps of the streaming COOP form with one "
                         "launch per pass (the headline's configuration)")
    ap.add_argument("--dim", type=int, default=N, help="grid edge (default: the 512^3 headline)")
    ap.add_argument("--radius", type=int, default=RADIUS)
    ap.add_argument("--sweeps", type=int, default=SWEEPS, help="sweeps per step (even)")
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--init", default="seeded", choices=["seeded", "reference"],
                    help="seeded: the same boundary in both buffers; reference: u2 zero "
                         "(different shells: the streaming form's band fix-up runs after each pass)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU slab halo exchange is a "
                                   "design note only, so --gpus must be 1"}))
        sys.exit(2)

    import numpy as np

    import thesis_amd as st

    if st.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    if args.probe:
        probe(st, np)
        return
    if args.workload == "batched":
        out = batched_leg(st, np, max(1, min(args.steps, 5)))
        out["n_gpus"] = 1
        print(json.dumps(out), flush=True)
        return
    if args.workload == "mixed384":
        out = mixed_leg(st, np, max(1, min(args.steps, 5)), with_cpu=not args.no_cpu)
        out["n_gpus"] = 1
        print(json.dumps(out), flush=True)
        return

    if args.workload == "sizes":
        out = {"metric": "GCells/s at the reference's evaluation sizes", "unit": "GCells/s",
               "higher_is_better": True, "dtype": "f32", "n_gpus": 1,
               "legs": size_legs(st, np, max(1, min(args.steps, 5)))}
        print(json.dumps(out), flush=True)
        return

    out, headline = headline_leg(st, np, args)
    if headline and not args.no_legs:
        reps = max(1, min(args.steps, 3))
        out["legs"] = {"mixed384": mixed_leg(st, np, reps, with_cpu=not args.no_cpu),
                       "batched": batched_leg(st, np, reps),
                       "reference_sizes": size_legs(st, np, reps)}
    if headline and not args.no_warm:
        out["warm_start"] = warm_start_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
