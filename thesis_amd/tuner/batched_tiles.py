"""Tune the tile rules of the batched sweep (rows per wave and planes per
chunk) for classes of problems, with the variable-tree search and an
in-process evaluator that times graph-mode batched launches on the GPU.

    python -m thesis_amd.tuner.batched_tiles [--out thesis_amd/tuned/batched_tiles.json]

Problem classes: radius 1 or 2 x volume buckets (<= 32^3, <= 64^3, <= 128^3)
of the 512-problem uneven batch (thesis_amd.workloads.batched_shapes).  Each
class is tuned on its own problems; the result is the rule table that
stencil_dispatch_batched takes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .. import workloads
from .evaluate import FunctionEvaluator
from .optimise import TreeOptimiser
from .vartree import parse_vartree

BUCKETS = [32 ** 3, 64 ** 3, 128 ** 3]
HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_OUT = os.path.join(HERE, "tuned", "batched_tiles.json")


def classes(shapes):
    out = {}
    for s in shapes:
        nx, ny, nz, r = s
        cells = nx * ny * nz
        b = next(m for m in BUCKETS if cells <= m)
        out.setdefault((r, b), []).append(s)
    return out


def tune(n=512, seed=0x5EED0004, iters=10, repeat=3, out=DEFAULT_OUT, verbose=True):
    import thesis_amd as st

    shapes = workloads.batched_shapes(n, seed)
    stream = st.Stream()
    rules, report = [], []
    for (r, bucket), members in sorted(classes(shapes).items()):
        bufs = []
        probs = []
        for (nx, ny, nz, rr) in members:
            p = st.StencilParams.make(nx=nx, ny=ny, nz=nz, dimensions=3, radius=rr, op=st.OP_JACOBI)
            d1 = st.DeviceArray(p.shape, np.float32)
            d2 = st.DeviceArray(p.shape, np.float32)
            st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=seed, stream=stream)
            bufs += [d1, d2]
            probs.append(dict(nx=nx, ny=ny, nz=nz, radius=rr, iterations=iters, op=st.OP_JACOBI,
                              u1=d1, u2=d2))
        stream.sync()
        # plane-chunk form: TY = frame rows (0: as many as fit), ZCHUNK =
        # planes per tile (0: the whole z extent)
        values = {"TY": ["0", "4", "8", "16"], "ZCHUNK": ["16", "32", "64", "0"]}

        def score(v):
            rule = dict(radius=r, max_cells=bucket, coarsen_y=int(v["TY"]), zchunk=int(v["ZCHUNK"]))
            st.dispatch_batched(probs, [rule], st.BATCH_GRAPH, stream=stream)  # warm
            return st.dispatch_batched(probs, [rule], st.BATCH_GRAPH, stream=stream, timed=True)

        ev = FunctionEvaluator(score, repeat=repeat, aggregator=min)
        opt = TreeOptimiser(parse_vartree("{TY, ZCHUNK}"), values, ev, minimise=True)
        res = opt.optimise()
        for b in bufs:
            b.free()
        if res is None:
            continue
        val, ms = res
        rules.append(dict(radius=r, max_cells=bucket, coarsen_y=int(val["TY"]),
                          zchunk=int(val["ZCHUNK"])))
        cells = sum(a * b * c for a, b, c, _ in members)
        report.append(dict(radius=r, max_cells=bucket, problems=len(members), ms=ms,
                           gcells_per_s=cells * iters / ms / 1e6,
                           tests=[dict(rec.valuation, ms=rec.overall) for rec in ev.order]))
        if verbose:
            print(f"radius {r} <= {bucket} cells ({len(members)} problems): TY={val['TY']} "
                  f"ZCHUNK={val['ZCHUNK']}  {ms:.3f} ms", flush=True)
    doc = {"device": st.device_name(0), "date": time.strftime("%Y-%m-%d"),
           "workload": f"{n} uneven problems (seed {seed:#x}), {iters} sweeps per timing",
           "rules": rules, "report": report}
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    return rules


def load_rules(path=DEFAULT_OUT):
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f)["rules"]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--iters", type=int, default=10)
    a = ap.parse_args()
    sys.exit(0 if tune(iters=a.iters, out=a.out) else 1)
