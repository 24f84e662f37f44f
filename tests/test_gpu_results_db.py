"""The result databases filled on the GPU through the C ABI."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu


def test_evaluate_and_autotune_on_device(gpu, tmp_path):
    from thesis_amd import results_db as R
    print("[case] results_db evaluate 3D 64^3", flush=True)
    out = str(tmp_path / "results.json")
    R.evaluate_stencils(out, [3], {3: [64]}, ["base", "smem", "smem_padded"], [1, 2], [1, 2], [8],
                        host="box", repeat=3)
    db = json.load(open(out))
    for v in ("1_gpus_base", "1_gpus_smem_unroll_2", "1_gpus_smem_padded"):
        for r in ("1", "2"):
            t = db["3"]["64"][v][r]["8"]["box"]["heuristic"]["time"]
            assert len(t) == 3 and all(x > 0 for x in t), (v, r, t)
    print("[case] results_db autotune 3D 64^3", flush=True)
    tune = str(tmp_path / "results_autotune.json")
    values = {"BLOCK_X": ["32", "64"], "BLOCK_Y": ["1", "2", "4"], "BLOCK_Z": ["1", "2"]}
    R.autotune_stencils(tune, [3], {3: [64]}, ["smem"], [1], [1], iterations=8, repeat=3, values=values)
    tdb = json.load(open(tune))
    e = tdb["3"]["64"]["smem"]["1"]
    assert e["BLOCK_X"] * e["BLOCK_Y"] * e["BLOCK_Z"] in (64, 128, 256, 512)
    R.evaluate_stencils(out, [3], {3: [64]}, ["smem"], [1], [1], [8], host="box", repeat=2, tune_db=tune)
    db = json.load(open(out))
    assert len(db["3"]["64"]["1_gpus_smem"]["1"]["8"]["box"]["autotune"]["time"]) == 2


@pytest.mark.skipif(os.environ.get("THESIS_NO_PROFILE_TESTS") == "1",
                    reason="rocprofv3 counter passes switched off (THESIS_NO_PROFILE_TESTS=1)")
def test_rocprof_metrics_on_device(gpu, tmp_path):
    """results_db profile's counter passes (three rocprofv3 runs of the
    stencil CLI, each under its own time limit) mapped to the metric names
    of the reference's nvprof database entries
    (scripts/profile_configuration.sh:21, scripts/evaluate_stencils.py:118-171)."""
    from thesis_amd import results_db as R
    print("[case] rocprof metrics smem_register 128^3", flush=True)
    cmd = R.cli_command(3, 128, "smem_register", 1, 8)
    m = R.rocprof_metrics(cmd, str(tmp_path), timeout_s=90)
    for k in ("dram_read_throughput", "dram_write_throughput", "achieved_occupancy", "shared_ld_bank_conflict",
              "l2_tex_hit_rate"):
        assert k in m, (k, m)
    assert float(m["dram_read_throughput"].rstrip("GB/s")) > 0
    assert 0 < float(m["achieved_occupancy"]) <= 1
