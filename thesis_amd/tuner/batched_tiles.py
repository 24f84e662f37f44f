"""Tune the tile rules of the batched sweep (frame rows and planes per
chunk) for classes of problems, with the variable-tree search and an
in-process evaluator that times graph-mode batched launches on the GPU.

    python -m thesis_amd.tuner.batched_tiles [--out thesis_amd/tuned/batched_tiles.json]

Problem classes: radius 1 or 2 x volume buckets (<= 32^3, <= 64^3, <= 128^3)
of the 512-problem uneven batch (thesis_amd.workloads.batched_shapes).  The
search is by coordinates: each class's rule is tuned with the tree search
while the other classes keep their current rules, and every candidate is
scored on the whole batch (a class alone does not fill the GPU, which biases
its choice towards small tiles).  The result is the rule table that
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
# plane-chunk form: TY = frame rows (0: as many as fit), ZCHUNK = planes per
# tile (0: the whole z extent)
VALUES = {"TY": ["0", "4", "8", "16"], "ZCHUNK": ["8", "16", "32", "64", "0"]}


def classes(shapes):
    out = {}
    for s in shapes:
        nx, ny, nz, r = s
        cells = nx * ny * nz
        b = next(m for m in BUCKETS if cells <= m)
        out.setdefault((r, b), []).append(s)
    return out


def rule_table(rules: dict) -> list:
    """Rules in the order the library matches them: by radius, then by
    ascending max_cells (the first rule whose bound holds applies)."""
    return [rules[k] for k in sorted(rules)]


def tune(n=512, seed=0x5EED0004, iters=10, repeat=3, passes=2, out=DEFAULT_OUT, verbose=True):
    import thesis_amd as st

    shapes = workloads.batched_shapes(n, seed)
    stream = st.Stream()
    bufs, probs = [], []
    for i, (nx, ny, nz, r) in enumerate(shapes):
        p = st.StencilParams.make(nx=nx, ny=ny, nz=nz, dimensions=3, radius=r, op=st.OP_JACOBI)
        d1 = st.DeviceArray(p.shape, np.float32)
        d2 = st.DeviceArray(p.shape, np.float32)
        st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=seed + i, stream=stream)
        bufs += [d1, d2]
        probs.append(dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=iters, op=st.OP_JACOBI, u1=d1, u2=d2))
    stream.sync()
    groups = classes(shapes)
    rules = {k: dict(radius=k[0], max_cells=k[1], coarsen_y=0, zchunk=16) for k in groups}

    def batch_ms(table):
        st.dispatch_batched(probs, table, st.BATCH_GRAPH, stream=stream)  # warm
        return st.dispatch_batched(probs, table, st.BATCH_GRAPH, stream=stream, timed=True)

    report = {}
    best_ms = None
    for sweep in range(passes):
        for key in sorted(groups):
            def score(v, key=key):
                trial = dict(rules)
                trial[key] = dict(radius=key[0], max_cells=key[1], coarsen_y=int(v["TY"]), zchunk=int(v["ZCHUNK"]))
                return batch_ms(rule_table(trial))

            ev = FunctionEvaluator(score, repeat=repeat, aggregator=min)
            opt = TreeOptimiser(parse_vartree("{TY, ZCHUNK}"), VALUES, ev, minimise=True)
            res = opt.optimise()
            if res is None:
                continue
            val, ms = res
            rules[key] = dict(radius=key[0], max_cells=key[1], coarsen_y=int(val["TY"]), zchunk=int(val["ZCHUNK"]))
            best_ms = ms
            report[key] = dict(radius=key[0], max_cells=key[1], problems=len(groups[key]), pass_=sweep,
                               batch_ms=ms, tests=[dict(rec.valuation, ms=rec.overall) for rec in ev.order])
            if verbose:
                print(f"pass {sweep} radius {key[0]} <= {key[1]} cells ({len(groups[key])} problems): "
                      f"TY={val['TY']} ZCHUNK={val['ZCHUNK']}  batch {ms:.3f} ms", flush=True)
    for b in bufs:
        b.free()
    cells = workloads.batched_interior_cells(shapes)
    doc = {"device": st.device_name(0), "date": time.strftime("%Y-%m-%d"),
           "workload": f"{n} uneven problems (seed {seed:#x}), {iters} sweeps per timing, whole-batch scoring",
           "batch_ms": best_ms, "gcells_per_s": cells * iters / best_ms / 1e6 if best_ms else None,
           "rules": rule_table(rules),
           "report": [dict(v, **{"pass": v.pop("pass_")}) for _, v in sorted(report.items())]}
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    return doc["rules"]


def load_rules(path=DEFAULT_OUT):
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f)["rules"]


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--passes", type=int, default=2)
    a = ap.parse_args()
    sys.exit(0 if tune(iters=a.iters, passes=a.passes, out=a.out) else 1)
