"""Write the golden fixtures from the reference CPU oracle's own outputs.

Run ``oracle/ref_build.sh`` first (it compiles mrfarstad/thesis
src/stencil_cpu.cu for each configuration and writes raw float32 solution
files, include/stencil_utils.h:16-44, into oracle/_ref/solutions/).  This
script turns them into data fixtures:

* ``ref_thesis.json``  one record per configuration: knobs, byte length,
  SHA-256 of the float32 field, max |u|, and the sum in FP64;
* ``ref_thesis_small.npz``  the full fields of the configurations with at
  most 8000 cells (float32, x fastest).

The fixtures are data only (inputs are the knobs, outputs are the fields);
no reference source is copied.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SOL = os.path.join(ROOT, "oracle", "_ref", "solutions")


def main() -> int:
    if not os.path.isdir(SOL):
        print("run oracle/ref_build.sh first", file=sys.stderr)
        return 1
    records = []
    small = {}
    for name in sorted(os.listdir(SOL)):
        _, dim, it, r, d = name.split("_")
        dim, it, r, dims = int(dim), int(it), int(r), int(d[0])
        raw = open(os.path.join(SOL, name), "rb").read()
        u = np.frombuffer(raw, np.float32)
        assert u.size == dim ** dims, name
        key = f"dim{dim}_it{it}_r{r}_{dims}d"
        records.append({
            "key": key, "dim": dim, "dimensions": dims, "radius": r, "iterations": it,
            "op": "thesis", "init": "reference", "bytes": len(raw),
            "sha256": hashlib.sha256(raw).hexdigest(),
            "max_abs": float(np.max(np.abs(u))), "sum_f64": float(np.sum(u, dtype=np.float64)),
        })
        if u.size <= 8000:
            small[key] = u.copy()
    with open(os.path.join(HERE, "ref_thesis.json"), "w") as f:
        json.dump({"source": "mrfarstad/thesis src/stencil_cpu.cu via oracle/ref_build.sh",
                   "records": records}, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_thesis_small.npz"), **small)
    print(f"{len(records)} records, {len(small)} full fields")
    return 0


if __name__ == "__main__":
    sys.exit(main())
