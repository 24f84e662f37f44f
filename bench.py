"""Benchmark: 3D 7-point JACOBI sweep, 512^3 FP32, radius 1, 100 sweeps.

One "step" is one pass of the workload: 100 sweeps of the 512^3 grid, with
the field already resident in HBM (seeded boundary shell written on the GPU
once, zero interior; each step continues from the previous state).  Prints
one JSON line.

  value      GCells/s = 510^3 interior cells x 100 sweeps x steps / time
  roofline   dominant kernel vs the 8 TB/s HBM peak, algorithmic bytes
             4 B x (512^3 + 510^3) per sweep, launch duration from HIP events
             on the launch stream
  cpu_baseline  the CPU oracle (C, OpenMP) on a bounded sample of the same
             workload, on this host's cores
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
    (the tuned U-Net's 64^3 prediction, upsampled) against the plain zero
    start, and the same with 2000 coarse JACOBI sweeps before the upsampling
    (and that refinement of the network's input guess alone, the control).
    Two sets of `problems` each: the five the network's training was judged
    on, and five further ones never used in any choice."""
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
                         "learned_ratio": round(r["time_ratio_learned_vs_plain"], 4),
                         "learned_coarse_ratio": round(r["time_ratio_learned_coarse_vs_plain"], 4),
                         "blend_coarse_ratio": round(r["time_ratio_blend_coarse_vs_plain"], 4)})
    mean = lambda k, sel=None: round(float(np.mean([r[k] for r in rows if sel is None or r["set"] == sel])), 4)
    return {"metric": "time to relative residual 1e-6 at 128^3: learned warm start / plain zero start",
            "unit": "ratio", "higher_is_better": False, "problems": len(rows),
            "learned": mean("learned_ratio"), "learned_standard": mean("learned_ratio", "standard"),
            "learned_further": mean("learned_ratio", "further"),
            "learned_max": round(max(r["learned_ratio"] for r in rows), 4),
            "learned_coarse": mean("learned_coarse_ratio"), "blend_coarse": mean("blend_coarse_ratio"),
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
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def bench_batched(st, np, args):
    """512 uneven 3D grids (edges log-uniform in [16, 128], radius 1 or 2),
    50 JACOBI sweeps each: one hipGraph launch, one persistent cooperative
    launch, and per-problem launches."""
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
    cells = workloads.batched_interior_cells(shapes) * its
    traffic = 4 * (workloads.batched_cells(shapes) + workloads.batched_interior_cells(shapes)) * its
    res = {}
    for name, mode in (("graph", st.BATCH_GRAPH), ("persistent", st.BATCH_PERSISTENT),
                       ("sequential", st.BATCH_SEQUENTIAL)):
        st.dispatch_batched(probs, rules, mode, stream=stream)  # warm-up
        t = min(st.dispatch_batched(probs, rules, mode, stream=stream, timed=True)
                for _ in range(max(1, args.steps)))
        res[name] = {"ms": round(t, 3), "gcells_per_s": round(cells / t / 1e6, 2),
                     "hbm_frac": round(traffic / (t / 1e3) / 8e12, 4)}
    out = {"metric": "batched GCells/s (512 uneven grids, 50 sweeps)", "value": res["graph"]["gcells_per_s"],
           "unit": "GCells/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic: 512 grids, edges round(16*8^U) per axis, radius 1|2, seeded boundaries",
           "config": {"workload": "batched uneven 3D JACOBI, 50 sweeps", "total_cells": workloads.batched_cells(shapes),
                      "tile_rules": rules},
           "modes": res,
           "speedup_graph_vs_sequential": round(res["sequential"]["ms"] / res["graph"]["ms"], 2),
           "speedup_persistent_vs_sequential": round(res["sequential"]["ms"] / res["persistent"]["ms"], 2)}
    print(json.dumps(out), flush=True)


def bench_mixed(st, np, args):
    """384^3, radius 4 (25-point), JACOBI c_r = 1/24, 200 sweeps: FP32, FP64 and
    BF16/FP16 storage with FP32 defect correction every 10 sweeps.  Errors are
    measured against the FP64 field on the GPU (whose bit-exactness with the CPU
    FP64 oracle the parity tests establish)."""
    n, r, its = 384, 4, 200
    seed = 0x5EED0003
    base = dict(dim=n, radius=r, iterations=its, op=st.OP_JACOBI, variant="auto", heuristic=1)
    shape = (n, n, n)
    interior = (n - 2 * r) ** 3
    out = {"metric": "384^3 R=4 error vs FP64 oracle per storage type", "unit": "GCells/s",
           "config": {"workload": "3D 25-point JACOBI, 384^3, 200 sweeps, restart 10"}, "results": {}}
    p64 = st.StencilParams.make(dtype=st.F64, **base)
    a = st.DeviceArray(shape, np.float64)
    b = st.DeviceArray(shape, np.float64)
    st.initialize_region(p64, a, b, st.INIT_SEEDED, seed=seed)
    _, ms = st.dispatch_kernels(p64, a, b, timed=True)
    ref = a.numpy()
    a.free(); b.free()
    out["results"]["f64"] = {"ms": round(ms, 3), "gcells_per_s": round(interior * its / ms / 1e6, 1),
                             "max_abs": 0.0, "l2": 0.0}

    def errs(u):
        d = u.astype(np.float64) - ref
        return float(np.max(np.abs(d))), float(np.sqrt(np.sum(d * d)) / np.sqrt(np.sum(ref * ref)))

    p32 = st.StencilParams.make(dtype=st.F32, **base)
    a = st.DeviceArray(shape, np.float32)
    b = st.DeviceArray(shape, np.float32)
    st.initialize_region(p32, a, b, st.INIT_SEEDED, seed=seed)
    x0 = a.numpy()
    _, ms = st.dispatch_kernels(p32, a, b, timed=True)
    mx, l2 = errs(a.numpy())
    out["results"]["f32"] = {"ms": round(ms, 3), "gcells_per_s": round(interior * its / ms / 1e6, 1),
                             "max_abs": mx, "l2_rel": l2}
    for name, dt in (("bf16", st.BF16), ("f16", st.F16)):
        pm = st.StencilParams.make(dtype=dt, restart=10, **{k: v for k, v in base.items() if k != "variant"})
        x = st.DeviceArray.from_numpy(x0)
        e1 = st.DeviceArray(shape, np.uint16)
        e2 = st.DeviceArray(shape, np.uint16)
        d = st.DeviceArray(shape, np.uint16)
        u = st.DeviceArray(shape, np.float32)
        ms = st.dispatch_mixed(pm, x, e1, e2, d, u, timed=True)
        mx, l2 = errs(u.numpy())
        out["results"][name] = {"ms": round(ms, 3), "gcells_per_s": round(interior * its / ms / 1e6, 1),
                                "max_abs": mx, "l2_rel": l2}
        for t in (x, e1, e2, d, u):
            t.free()
    out["value"] = out["results"]["f32"]["gcells_per_s"]
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--workload", default="sweep512",
                    choices=["sweep512", "batched", "mixed384"],
                    help="sweep512: the headline 512^3 sweep (default); batched: 512 uneven "
                         "grids in one graph / persistent launch vs per-problem launches; "
                         "mixed384: 384^3 R=4 in FP32, FP64 and 16-bit storage with errors "
                         "against the FP64 oracle")
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
                    help="seeded: the same boundary in both buffers (two sweeps fuse per pass "
                         "in the streaming form); reference: u2 zero (different shells, one "
                         "sweep per pass)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU slab halo exchange "
                                   "is a design note only, so --gpus must be 1"}))
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
        bench_batched(st, np, args)
        return
    if args.workload == "mixed384":
        bench_mixed(st, np, args)
        return

    headline = (args.dim, args.radius, args.sweeps, args.dtype) == (N, RADIUS, SWEEPS, "f32")
    # the committed traffic profiles were measured on the seeded field (equal
    # shells: fused sweeps, whole-row stores)
    measured_traffic = headline and args.init == "seeded"
    if args.sweeps % 2:
        raise SystemExit("--sweeps must be even (the state stays in the first buffer)")
    N, RADIUS, SWEEPS = args.dim, args.radius, args.sweeps
    f64 = args.dtype == "f64"
    es = 8 if f64 else 4
    kw = dict(dim=N, radius=RADIUS, iterations=SWEEPS, op=st.OP_JACOBI, dtype=st.F64 if f64 else st.F32,
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

    # The field is initialised once; every step advances it by 100 more
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
    cells = (N - 2 * RADIUS) ** 3 * SWEEPS * args.steps
    gcells = cells / wall / 1e9
    launches = cfg["launches"]
    bytes_sweep = algorithmic_bytes_per_sweep(N, RADIUS, es)
    bytes_launch = bytes_sweep * SWEEPS / launches
    avg_launch_s = sweep_total_s / (launches * args.steps)
    achieved = bytes_launch / avg_launch_s / 1e9
    variant_name = {v: k for k, v in st.VARIANTS.items()}[cfg["variant"]]
    out = {
        "metric": "GCells/s & %HBM roofline (512^3 r=1)" if headline
                  else f"GCells/s & %HBM roofline ({N}^3 r={RADIUS} {args.dtype})",
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
        "config": {"workload": (f"3D {6 * RADIUS + 1}-point JACOBI u'=(1/{6 * RADIUS})*sum of "
                                f"{6 * RADIUS} neighbours, {N}^3 {args.dtype.upper()}, radius {RADIUS}, "
                                f"{SWEEPS} sweeps per step"),
                   "variant": variant_name, "launch": cfg, "device": st.device_name(0)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(variant_name) if measured_traffic else None,
                     "bytes_per_launch_algorithmic": bytes_launch,
                     "avg_launch_ms": round(avg_launch_s * 1e3, 5),
                     "note": ("temporal-blocked passes advance two sweeps per HBM pass, so the "
                              "algorithmic fraction can exceed 1; a launch is one pass, and the "
                              "event time per launch includes the shell comparison and the band "
                              "fix-up launches (which return at once with equal shells)")
                             if variant_name == "coop"
                             else "one sweep per launch"},
    }
    d1.free()
    d2.free()
    if headline and not args.no_warm:
        out["warm_start"] = warm_start_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
