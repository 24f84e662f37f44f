"""The result databases keep the reference's key layout and resume rules
(scripts/evaluate_stencils.py:40-109,227-231, scripts/autotune_configuration.py:27-79).
A fake timer stands in for the GPU."""
import json
from functools import reduce

import pytest

from thesis_amd import results_db as R


def dotted_get(dictionary, keys, default=None):
    # the reference's own lookup (evaluate_stencils.py:40-45)
    return reduce(lambda d, key: d.get(key, default) if isinstance(d, dict) else default,
                  keys.split("."), dictionary)


def test_version_keys():
    assert R.version_key("base", 1) == "1_gpus_base"
    assert R.version_key("smem", 4) == "1_gpus_smem_unroll_4"
    assert R.version_key("smem_padded", 2, gpus=8) == "8_gpus_smem_padded_unroll_2"
    assert R.tune_key("smem", 1) == "smem" and R.tune_key("smem", 8) == "smem_unroll_8"


def test_evaluate_layout_and_resume(tmp_path):
    out = str(tmp_path / "results.json")
    calls = []

    def timer(p, repeat):
        calls.append((p.variant, p.radius, p.coarsen_x, p.iterations, p.heuristic, repeat))
        if p.radius == 16:
            raise RuntimeError("rejected")
        return [1.0 + p.radius + 0.1 * i for i in range(repeat)]

    R.evaluate_stencils(out, [3], {3: [64]}, ["base", "smem"], [1, 2], [1, 16], [8], host="box",
                        repeat=3, timer=timer)
    db = json.load(open(out))
    assert dotted_get(db, "3.64.1_gpus_base.1.8.box.heuristic.time") == [2.0, 2.1, 2.2]
    assert dotted_get(db, "3.64.1_gpus_smem_unroll_2.1.8.box.heuristic.time") == [2.0, 2.1, 2.2]
    assert dotted_get(db, "3.64.1_gpus_base.16.8.box.heuristic.time") == []   # failed run
    assert len(calls) == 8 and all(c[4] == 1 and c[5] == 3 for c in calls)
    assert {c[2] for c in calls} == {1, 2}

    def no_timer(p, repeat):
        raise AssertionError("existing entries must not be measured again")

    R.evaluate_stencils(out, [3], {3: [64]}, ["base", "smem"], [1, 2], [1, 16], [8], host="box",
                        repeat=3, timer=no_timer)
    # a new iteration count adds entries beside the old ones
    R.evaluate_stencils(out, [3], {3: [64]}, ["base"], [1], [1], [16], host="box", repeat=2,
                        timer=timer)
    db = json.load(open(out))
    assert dotted_get(db, "3.64.1_gpus_base.1.16.box.heuristic.time") == [2.0, 2.1]
    assert dotted_get(db, "3.64.1_gpus_base.1.8.box.heuristic.time") == [2.0, 2.1, 2.2]


def test_autotune_then_evaluate_with_tuned_blocks(tmp_path):
    tune = str(tmp_path / "results_autotune.json")
    best = (64, 4, 2)

    def timer(p, repeat):
        b = (p.block_x, p.block_y, p.block_z)
        return [10.0 + sum(abs(x - y) for x, y in zip(b, best))] * repeat

    values = {"BLOCK_X": ["16", "32", "64", "128"], "BLOCK_Y": ["1", "2", "4", "8"],
              "BLOCK_Z": ["1", "2", "4"]}
    R.autotune_stencils(tune, [3], {3: [128]}, ["smem"], [1, 2], [1], iterations=8, repeat=3,
                        timer=timer, values=values, csv_dir=str(tmp_path / "csv"))
    tdb = json.load(open(tune))
    assert tdb["3"]["128"]["smem"]["1"] == {"BLOCK_X": 64, "BLOCK_Y": 4, "BLOCK_Z": 2}
    assert tdb["3"]["128"]["smem_unroll_2"]["1"] == {"BLOCK_X": 64, "BLOCK_Y": 4, "BLOCK_Z": 2}
    assert (tmp_path / "csv" / "tune_3d_128_smem_r1.csv").exists()

    seen = []

    def record(p, repeat):
        seen.append((p.heuristic, p.block_x, p.block_y, p.block_z))
        return [1.0] * repeat

    out = str(tmp_path / "results.json")
    R.evaluate_stencils(out, [3], {3: [128]}, ["smem"], [1, 2, 4], [1], [8], host="box", repeat=2,
                        tune_db=tune, timer=record)
    db = json.load(open(out))
    assert dotted_get(db, "3.128.1_gpus_smem.1.8.box.autotune.time") == [1.0, 1.0]
    # no tuned shape for unroll 4: skipped, as the reference does
    assert dotted_get(db, "3.128.1_gpus_smem_unroll_4.1.8.box.autotune") is None
    assert seen == [(0, 64, 4, 2), (0, 64, 4, 2)]


def test_params_follow_the_reference_knobs():
    p = R.make_params(3, 64, "smem_padded", 2, 8, unroll=4, block=(32, 4, 2), smem_pad=1)
    assert (p.nx, p.ny, p.nz, p.radius, p.iterations, p.coarsen_x, p.smem_pad) == (64, 64, 64, 2, 8, 4, 1)
    assert (p.heuristic, p.block_x, p.block_y, p.block_z) == (0, 32, 4, 2)
    q = R.make_params(2, 256, "base", 1, 64)
    assert (q.nx, q.ny, q.nz, q.dimensions, q.heuristic) == (256, 256, 1, 2, 1)


def test_profile_layout_and_skip(tmp_path):
    out = str(tmp_path / "results.json")
    seen = []

    def metrics(cmd):
        seen.append(cmd)
        return {"dram_read_throughput": "123.4000GB/s", "shared_ld_bank_conflict": "0.5000%",
                "achieved_occupancy": "0.812500"}

    R.profile_stencils(out, [3], {3: [64]}, ["base", "smem"], [1], [1, 2], [8], host="box", metrics=metrics)
    db = json.load(open(out))
    e = dotted_get(db, "3.64.1_gpus_smem.2.8.box.heuristic")
    assert e["dram_read_throughput"] == "123.4000GB/s" and e["achieved_occupancy"] == "0.812500"
    assert len(seen) == 4
    cmd = seen[0]
    assert cmd[0].endswith("stencil") and cmd[cmd.index("--variant") + 1] == "base"
    assert cmd[cmd.index("--radius") + 1] == "1" and cmd[cmd.index("--heuristic") + 1] == "1"
    # entries with metrics are not profiled again; "time" beside them survives
    R.evaluate_stencils(out, [3], {3: [64]}, ["base"], [1], [1], [8], host="box", repeat=2,
                        timer=lambda p, rep: [1.0] * rep)
    R.profile_stencils(out, [3], {3: [64]}, ["base", "smem"], [1], [1, 2], [8], host="box",
                       metrics=lambda cmd: (_ for _ in ()).throw(AssertionError("profiled twice")))
    db = json.load(open(out))
    e = dotted_get(db, "3.64.1_gpus_base.1.8.box.heuristic")
    assert e["time"] == [1.0, 1.0] and e["shared_ld_bank_conflict"] == "0.5000%"


def test_read_counters_skips_helper_kernels(tmp_path):
    p = tmp_path / "p_counter_collection.csv"
    p.write_text(
        '"Correlation_Id","Dispatch_Id","Kernel_Name","Scratch_Size","Counter_Name","Counter_Value",'
        '"Start_Timestamp","End_Timestamp"\n'
        '1,1,"void st::k_init<float>(float*)",0,"FETCH_SIZE",10.0,1000,2000\n'
        '2,2,"void st::k_rows<1, 1, 4, float, 0, false>(void const*)",0,"FETCH_SIZE",1024.0,0,1000000\n'
        '3,3,"void st::k_rows<1, 1, 4, float, 0, false>(void const*)",0,"FETCH_SIZE",2048.0,0,2000000\n')
    got = R._read_counters(str(p))
    assert [v for v, _, _ in got["FETCH_SIZE"]] == [1024.0, 2048.0]
    assert got["FETCH_SIZE"][1][1] == pytest.approx(2e-3)
