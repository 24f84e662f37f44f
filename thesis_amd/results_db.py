"""Result databases in the reference's JSON wire formats, filled from the C ABI.

Two files, as the reference's experiment drivers keep them:

* the measurement database (``scripts/evaluate_stencils.py:60-109,227-231``)::

      db[DIMENSIONS][DIM][VERSION][RADIUS][ITERATIONS][HOST][CONFIG]["time"]
          = [ms, ms, ...]                       # one entry per repeat

  with ``VERSION = "<gpus>_gpus_<version>[_unroll_<COARSEN_X>]"``
  (``:80-84``), ``CONFIG`` ``"heuristic"`` or ``"autotune"`` (``:28``),
  every key a string, and the metrics of a profiled configuration stored
  beside ``"time"`` under their own names (``:170-173``);
* the autotune database (``scripts/autotune_configuration.py:27-79``)::

      tune_db[DIMENSIONS][DIM][version[_unroll_<COARSEN_X>]][RADIUS]
          = {"BLOCK_X": int, "BLOCK_Y": int, "BLOCK_Z": int}

Both drivers are resumable in the reference's way: an entry that already
exists is not measured again (``evaluate_stencils.py:47-58,175-189``), and
the file is rewritten after every new entry.  The reference runs a
per-configuration binary ``REPEAT`` times and parses each run's last line
(``scripts/evaluate_configuration.sh:40``); here each repeat initialises the
two buffers on the GPU (the reference initial condition, buffer 2 zero) and
times one ``stencil_dispatch`` of ``ITERATIONS`` sweeps with the library's
stream events, which is the number the ``stencil`` CLI prints.  The search
space of the autotuner is the reference's (powers of two 1..1024 per block
axis, ``scripts/autotune_configuration.sh:17-24``) minus the shapes a 64-lane
wave cannot run (``tuner.optimise.wave64_block_filter``).

Command line::

    python -m thesis_amd.results_db evaluate --out results/results_stencils.json \\
        --dimensions 3 --dims 256 --versions base smem smem_padded --depths 1 2 4 \\
        --unrolls 1 2 --iterations 8 --repeat 30 [--autotune results/results_autotune.json]
    python -m thesis_amd.results_db autotune --out results/results_autotune.json ...
    python -m thesis_amd.results_db profile --out results/results_stencils.json ...

The profile mode adds, beside "time", the hardware metrics the reference
collected with nvprof (scripts/profile_configuration.sh), under the same
names, from rocprofv3 counters (see rocprof_metrics for the mapping).
"""
from __future__ import annotations

import json
import os
import socket
import statistics
from typing import Callable, Dict, Iterable, List, Optional

VERSIONS = ("base", "smem", "smem_padded", "smem_register", "coop")

# The reference's fixed block shape for heuristic runs
# (scripts/evaluate_stencils.py:108-110); with HEURISTIC=1 the library picks
# its own shape and these are informational only.
HEURISTIC_BLOCK = {"2": (32, 32, 1), "3": (32, 8, 4)}


def version_key(version: str, unroll: int, gpus: int = 1) -> str:
    """``1_gpus_smem_unroll_4`` (scripts/evaluate_stencils.py:80-84)."""
    v = (f"{gpus}_gpus_" if gpus > 0 else "_gpu_") + version
    if unroll > 1:
        v += f"_unroll_{unroll}"
    return v


def tune_key(version: str, unroll: int) -> str:
    """``smem_unroll_4`` (scripts/autotune_configuration.py:42-45)."""
    return version + (f"_unroll_{unroll}" if unroll > 1 else "")


def deep_get(d, keys: Iterable, default=None):
    """The reference's dotted lookup (evaluate_stencils.py:40-45) on a list of keys."""
    for k in keys:
        if not isinstance(d, dict):
            return default
        d = d.get(str(k), default)
    return d


def _load(path: Optional[str]) -> dict:
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _save(db: dict, path: Optional[str]):
    if not path:
        return
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f)
    os.replace(tmp, path)


def _node(db: dict, keys: List) -> dict:
    for k in keys:
        db = db.setdefault(str(k), {})
    return db


# ---------------------------------------------------------------- measuring

def make_params(dimension: int, dim: int, version: str, radius: int, iterations: int,
                unroll: int = 1, block=None, smem_pad: int = 0):
    """st_params of one reference configuration (the knobs
    scripts/set_run_configuration.sh writes into constants.sh).  block=None
    means HEURISTIC=1."""
    import thesis_amd as st

    kw = dict(dim=dim, dimensions=dimension, radius=radius, iterations=iterations,
              variant=version, coarsen_x=unroll, smem_pad=smem_pad, op=st.OP_THESIS)
    if dimension == 2:
        kw["nz"] = 1
    if block is None:
        kw["heuristic"] = 1
    else:
        bx, by, bz = block
        kw.update(heuristic=0, block_x=bx, block_y=by, block_z=bz if dimension == 3 else 1)
    return st.StencilParams.make(**kw)


class DeviceRunner:
    """Times configurations on the GPU through the C ABI.  One pair of
    buffers is kept per shape."""

    def __init__(self):
        import numpy as np

        import thesis_amd as st

        self.st, self.np = st, np
        self.stream = st.Stream()
        self._bufs = {}

    def _buffers(self, p):
        key = tuple(p.shape)
        if key not in self._bufs:
            for b in self._bufs.values():
                b[0].free()
                b[1].free()
            self._bufs = {key: (self.st.DeviceArray(p.shape, self.np.float32),
                                self.st.DeviceArray(p.shape, self.np.float32))}
        return self._bufs[key]

    def time(self, p, repeat: int) -> List[float]:
        """`repeat` runs: initialise (reference initial condition), then one
        timed dispatch of p.iterations sweeps.  Raises StencilError for a
        configuration the library rejects (the reference's exit code 1)."""
        st = self.st
        st.check_early_exit(p)
        d1, d2 = self._buffers(p)
        out = []
        for _ in range(repeat):
            st.initialize_region(p, d1, d2, st.INIT_REFERENCE, stream=self.stream)
            _, ms = st.dispatch_kernels(p, d1, d2, stream=self.stream, timed=True)
            out.append(float(ms))
        return out

    def close(self):
        for b in self._bufs.values():
            b[0].free()
            b[1].free()
        self._bufs = {}


def evaluate_stencils(out: str, dimensions: Iterable[int] = (3,), dims: Optional[Dict[int, List[int]]] = None,
                      versions: Iterable[str] = ("base", "smem", "smem_padded"),
                      unrolls: Iterable[int] = (1, 2, 4, 8), depths: Iterable[int] = (1, 2, 4, 8, 16),
                      iterations: Iterable[int] = (8,), host: Optional[str] = None, repeat: int = 30,
                      gpus: int = 1, tune_db: Optional[str] = None,
                      timer: Optional[Callable] = None, verbose: bool = False) -> dict:
    """Fill the measurement database `out` (created or resumed).  `timer(p,
    repeat) -> [ms]` defaults to DeviceRunner().time; a configuration it
    rejects (StencilError / ValueError) gets an empty time list, as a failed
    reference run yields no numbers."""
    db = _load(out)
    tdb = _load(tune_db) if tune_db else {}
    autotune = bool(tune_db)
    config = "autotune" if autotune else "heuristic"
    host = host or socket.gethostname().split(".")[0]
    dims = dims or {3: [256, 1024], 2: [4096, 32768]}
    runner = None
    if timer is None:
        runner = DeviceRunner()
        timer = runner.time
    try:
        for dimension in dimensions:
            for dim in dims[dimension]:
                for version in versions:
                    for unroll in unrolls:
                        v = version_key(version, unroll, gpus)
                        for depth in depths:
                            for it in iterations:
                                keys = [dimension, dim, v, depth, it, host, config]
                                if deep_get(db, keys + ["time"]) is not None:
                                    continue
                                block = None
                                if autotune:
                                    bd = deep_get(tdb, [dimension, dim, tune_key(version, unroll), depth])
                                    if not bd:
                                        continue  # the reference falls back per entry too
                                    block = (int(bd["BLOCK_X"]), int(bd["BLOCK_Y"]),
                                             int(bd.get("BLOCK_Z", 1)))
                                p = make_params(dimension, dim, version, depth, it, unroll, block)
                                try:
                                    times = list(timer(p, repeat))
                                except (ValueError, RuntimeError) as e:
                                    if verbose:
                                        print(f"[results_db] {keys}: rejected ({e})", flush=True)
                                    times = []
                                _node(db, keys)["time"] = times
                                _save(db, out)
                                if verbose and times:
                                    print(f"[results_db] {keys}: median {statistics.median(times):.4f} ms",
                                          flush=True)
    finally:
        if runner is not None:
            runner.close()
    _save(db, out)
    return db


POW2 = [str(1 << i) for i in range(11)]  # 1 .. 1024 (scripts/autotune_configuration.sh:17-24)


def autotune_stencils(out: str, dimensions: Iterable[int] = (3,), dims: Optional[Dict[int, List[int]]] = None,
                      versions: Iterable[str] = ("base", "smem", "smem_padded"),
                      unrolls: Iterable[int] = (1, 2, 4, 8), depths: Iterable[int] = (1, 2, 4, 8, 16),
                      iterations: int = 8, repeat: int = 30, csv_dir: Optional[str] = None,
                      timer: Optional[Callable] = None, values: Optional[Dict[str, List[str]]] = None,
                      verbose: bool = False) -> dict:
    """Fill the autotune database `out`: for every (dimension, dim,
    version[_unroll], radius) the block shape with the smallest median time
    over `repeat` runs (the tuner's `repeat = N, median`), searched with the
    variable tree {BLOCK_X, BLOCK_Y, BLOCK_Z}."""
    from .tuner.evaluate import FunctionEvaluator
    from .tuner.optimise import TreeOptimiser, wave64_block_filter
    from .tuner.vartree import parse_vartree

    db = _load(out)
    dims = dims or {3: [1024], 2: [32768]}
    runner = None
    if timer is None:
        runner = DeviceRunner()
        timer = runner.time
    try:
        for dimension in dimensions:
            names = ["BLOCK_X", "BLOCK_Y", "BLOCK_Z"] if dimension == 3 else ["BLOCK_X", "BLOCK_Y"]
            space = values or {n: POW2 for n in names}
            if dimension == 2:
                space = dict(space, BLOCK_Z=["1"])
            for dim in dims[dimension]:
                for version in versions:
                    for unroll in unrolls:
                        v = tune_key(version, unroll)
                        for depth in depths:
                            keys = [dimension, dim, v, depth]
                            if deep_get(db, keys) is not None:
                                continue

                            def score(val, _d=dimension, _n=dim, _v=version, _r=depth, _u=unroll):
                                p = make_params(_d, _n, _v, _r, iterations, _u,
                                                (int(val["BLOCK_X"]), int(val["BLOCK_Y"]),
                                                 int(val["BLOCK_Z"])))
                                return statistics.median(timer(p, repeat))

                            ev = FunctionEvaluator(score, repeat=1)
                            tree = parse_vartree("{BLOCK_X, BLOCK_Y, BLOCK_Z}")
                            res = TreeOptimiser(tree, space, ev, minimise=True,
                                                accept=wave64_block_filter if dimension == 3 else
                                                _wave64_2d).optimise()
                            if csv_dir:
                                os.makedirs(csv_dir, exist_ok=True)
                                ev.write_csv(os.path.join(csv_dir, f"tune_{dimension}d_{dim}_{v}_r{depth}.csv"),
                                             ["BLOCK_X", "BLOCK_Y", "BLOCK_Z"])
                            if res is None:
                                continue
                            val, ms = res
                            entry = {n: int(val[n]) for n in names}
                            _node(db, keys[:-1])[str(depth)] = entry
                            _save(db, out)
                            if verbose:
                                print(f"[autotune] {keys}: {entry} {ms:.4f} ms", flush=True)
    finally:
        if runner is not None:
            runner.close()
    _save(db, out)
    return db


# ---------------------------------------------------------------- profiling

# rocprofv3 counter passes (one run each: rocprofv3 does not split a request
# over passes) and the calibration of the memory counters against the
# library's streaming copy kernel (profiles/traffic_512cubed_r1.md: FETCH_SIZE
# counts half of the bytes read on this GPU, WRITE_SIZE all of them).
PROFILE_PASSES = (("FETCH_SIZE",), ("WRITE_SIZE",),
                  ("OccupancyPercent", "LDSBankConflict", "VALUBusy", "TCC_HIT_sum", "TCC_MISS_sum"))
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
_HELPER_KERNELS = ("k_init", "k_shell_eq", "k_copy_f32", "k_fill16")


def cli_command(dimension: int, dim: int, version: str, radius: int, iterations: int, unroll: int = 1,
                block=None, form: int = 0) -> List[str]:
    """The `stencil` CLI line of one reference configuration (the knobs of
    scripts/set_run_configuration.sh and profile_configuration.sh)."""
    here = os.path.dirname(os.path.abspath(__file__))
    cmd = [os.path.join(here, "bin", "stencil"), "--dim", str(dim), "--dimensions", str(dimension),
           "--radius", str(radius), "--iterations", str(iterations), "--variant", version,
           "--coarsen-x", str(unroll)]
    if form:
        cmd += ["--form", str(form)]
    if block is None:
        cmd += ["--heuristic", "1"]
    else:
        cmd += ["--heuristic", "0", "--block-x", str(block[0]), "--block-y", str(block[1]),
                "--block-z", str(block[2] if dimension == 3 else 1)]
    return cmd


def _read_counters(csv_path: str):
    """{counter: [(value, duration_s, scratch_bytes)] over the sweep kernels'
    dispatches} from a rocprofv3 counter_collection.csv."""
    import csv

    out: Dict[str, list] = {}
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name", "")
            if any(h in name for h in _HELPER_KERNELS):
                continue
            dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
            out.setdefault(r["Counter_Name"], []).append(
                (float(r["Counter_Value"]), dur, int(float(r.get("Scratch_Size", 0) or 0))))
    return out


def rocprof_metrics(cmd: List[str], workdir: str, timeout_s: int = 120) -> Dict[str, str]:
    """Run `cmd` under rocprofv3 once per counter pass and map the counters
    to the metric names of the reference's nvprof database entries
    (scripts/profile_configuration.sh, read by scripts/plot_measurements.py,
    which keeps only the number of each string):

      dram_read_throughput / dram_write_throughput   calibrated FETCH_SIZE /
                                                     WRITE_SIZE per second
      achieved_occupancy      OccupancyPercent / 100 (fraction of 32 waves/CU)
      shared_ld_bank_conflict LDSBankConflict: % of GPU time LDS stalled on
                              bank conflicts (nvprof counted conflicts)
      l2_tex_hit_rate         100 * TCC_HIT / (TCC_HIT + TCC_MISS) (the
                              derived L2CacheHit has no gfx950 definition)
      local_load_throughput, local_store_throughput
                              0 when the kernels use no scratch (spill) memory
    plus AMD-named values (valu_busy_pct, kernel_ms).  Counter values are
    averaged over the sweep kernels' dispatches."""
    import glob
    import shutil
    import signal
    import subprocess

    env = dict(os.environ, TMPDIR="/tmp")
    got: Dict[str, list] = {}
    for i, counters in enumerate(PROFILE_PASSES):
        d = os.path.join(workdir, f"pass{i}")
        shutil.rmtree(d, ignore_errors=True)
        prof = ["rocprofv3", "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "p", "--", *cmd]
        # its own process group, so that a pass over its time limit is ended
        # with the program it profiles
        proc = subprocess.Popen(prof, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                start_new_session=True)
        try:
            stdout, _ = proc.communicate(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.communicate()
            raise RuntimeError(f"rocprofv3 pass {counters} exceeded {timeout_s} s")
        files = glob.glob(os.path.join(d, "**", "p_counter_collection.csv"), recursive=True)
        if not files:
            raise RuntimeError(f"rocprofv3 pass {counters} wrote no counters (exit {proc.returncode}): "
                               f"{stdout.decode(errors='replace')[-400:]}")
        got.update(_read_counters(files[0]))

    def mean(name, idx=0):
        v = got.get(name)
        if not v:
            return None
        return sum(x[idx] for x in v) / len(v)

    out: Dict[str, str] = {}
    fetch, write = got.get("FETCH_SIZE"), got.get("WRITE_SIZE")
    if fetch:
        gbs = sum(v * 1024.0 * FETCH_SCALE / t for v, t, _ in fetch) / len(fetch) / 1e9
        out["dram_read_throughput"] = f"{gbs:.4f}GB/s"
        out["kernel_ms"] = f"{1e3 * sum(t for _, t, _ in fetch) / len(fetch):.6f}"
    if write:
        gbs = sum(v * 1024.0 * WRITE_SCALE / t for v, t, _ in write) / len(write) / 1e9
        out["dram_write_throughput"] = f"{gbs:.4f}GB/s"
    occ = mean("OccupancyPercent")
    if occ is not None:
        out["achieved_occupancy"] = f"{occ / 100.0:.6f}"
    lds = mean("LDSBankConflict")
    if lds is not None:
        out["shared_ld_bank_conflict"] = f"{lds:.4f}%"
    hit, miss = mean("TCC_HIT_sum"), mean("TCC_MISS_sum")
    if hit is not None and miss is not None and hit + miss > 0:
        out["l2_tex_hit_rate"] = f"{100.0 * hit / (hit + miss):.2f}%"
    valu = mean("VALUBusy")
    if valu is not None:
        out["valu_busy_pct"] = f"{valu:.2f}%"
    scratch = [x[2] for v in got.values() for x in v]
    if scratch and max(scratch) == 0:
        out["local_load_throughput"] = "0.000000B/s"
        out["local_store_throughput"] = "0.000000B/s"
    return out


def profile_stencils(out: str, dimensions: Iterable[int] = (3,), dims: Optional[Dict[int, List[int]]] = None,
                     versions: Iterable[str] = ("base", "smem", "smem_padded"),
                     unrolls: Iterable[int] = (1,), depths: Iterable[int] = (1, 2, 4, 8, 16),
                     iterations: Iterable[int] = (8,), host: Optional[str] = None, gpus: int = 1,
                     tune_db: Optional[str] = None, workdir: str = "/tmp/thesis_amd_profile",
                     metrics: Optional[Callable] = None, verbose: bool = False) -> dict:
    """Fill the metric entries of the measurement database `out` beside
    "time" (scripts/evaluate_stencils.py:112-173): an entry that already has
    "shared_ld_bank_conflict" is skipped, as the reference checks.
    `metrics(cmd) -> {name: str}` defaults to rocprof_metrics."""
    db = _load(out)
    tdb = _load(tune_db) if tune_db else {}
    config = "autotune" if tune_db else "heuristic"
    host = host or socket.gethostname().split(".")[0]
    dims = dims or {3: [256, 1024], 2: [4096, 32768]}
    metrics = metrics or (lambda cmd: rocprof_metrics(cmd, workdir))
    for dimension in dimensions:
        for dim in dims[dimension]:
            for version in versions:
                for unroll in unrolls:
                    v = version_key(version, unroll, gpus)
                    for depth in depths:
                        for it in iterations:
                            keys = [dimension, dim, v, depth, it, host, config]
                            if deep_get(db, keys + ["shared_ld_bank_conflict"]) is not None:
                                continue
                            block = None
                            if tune_db:
                                bd = deep_get(tdb, [dimension, dim, tune_key(version, unroll), depth])
                                if not bd:
                                    continue
                                block = (int(bd["BLOCK_X"]), int(bd["BLOCK_Y"]), int(bd.get("BLOCK_Z", 1)))
                            cmd = cli_command(dimension, dim, version, depth, it, unroll, block)
                            try:
                                vals = metrics(cmd)
                            except (RuntimeError, OSError, ValueError) as e:
                                if verbose:
                                    print(f"[results_db] {keys}: profile failed ({e})", flush=True)
                                continue
                            _node(db, keys).update(vals)
                            _save(db, out)
                            if verbose:
                                print(f"[results_db] {keys}: {vals}", flush=True)
    _save(db, out)
    return db


def _wave64_2d(valuation: Dict[str, str]) -> bool:
    bx, by = int(valuation["BLOCK_X"]), int(valuation["BLOCK_Y"])
    t = bx * by
    return t % 64 == 0 and 64 <= t <= 1024 and bx >= by


# ---------------------------------------------------------------- command line

def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(prog="python -m thesis_amd.results_db")
    ap.add_argument("mode", choices=["evaluate", "autotune", "profile"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--dimensions", type=int, nargs="+", default=[3])
    ap.add_argument("--dims", type=int, nargs="+", default=None,
                    help="grid edges (applied to every dimension given)")
    ap.add_argument("--versions", nargs="+", default=["base", "smem", "smem_padded"], choices=VERSIONS)
    ap.add_argument("--unrolls", type=int, nargs="+", default=[1, 2, 4, 8])
    ap.add_argument("--depths", type=int, nargs="+", default=[1, 2, 4, 8, 16])
    ap.add_argument("--iterations", type=int, nargs="+", default=[8])
    ap.add_argument("--repeat", type=int, default=30)
    ap.add_argument("--host", default=None)
    ap.add_argument("--autotune", default=None, help="autotune database to take block shapes from")
    ap.add_argument("--csv-dir", default=None, help="autotune: write the tuner's CSV log per entry")
    a = ap.parse_args(argv)
    dims = {d: a.dims for d in a.dimensions} if a.dims else None
    if a.mode == "evaluate":
        evaluate_stencils(a.out, a.dimensions, dims, a.versions, a.unrolls, a.depths, a.iterations,
                          a.host, a.repeat, tune_db=a.autotune, verbose=True)
    elif a.mode == "profile":
        profile_stencils(a.out, a.dimensions, dims, a.versions, a.unrolls, a.depths, a.iterations,
                         a.host, tune_db=a.autotune, verbose=True)
    else:
        autotune_stencils(a.out, a.dimensions, dims, a.versions, a.unrolls, a.depths, a.iterations[0],
                          a.repeat, csv_dir=a.csv_dir, verbose=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
