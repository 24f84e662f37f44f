"""The `stencil` CLI's process contract (src/main.cu:10-103), run as a process.

* the last stdout line is the kernel time in milliseconds, "%.4f"
  (src/main.cu:89), which the tuner and scripts/evaluate_configuration.sh:36
  parse;
* knobs that check_early_exit rejects exit with status 1
  (include/stencil_utils.h:113-124);
* --debug reads solutions/solution_<DIM>_<IT>_<R>_<D>d (raw float32, x
  fastest, include/stencil_utils.h:16-76), prints "reading ..." and then
  "rms error = %f" (include/stencil_error_checker.h:6-21), and exits 1 when
  the sum of squared differences exceeds 1e-5.  The solutions here are
  written by the CPU oracle, or are the reference's own stencil_cpu outputs
  (tests/golden/ref_thesis_small.npz), so every variant is checked against a
  CPU result the way scripts/test_all_configurations.sh:25-26 checks it
  ("rms error = 0.000000" on the line after "reading solution").
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "thesis_amd", "bin", "stencil")


def run_cli(args, cwd):
    print("[cli]", " ".join(str(a) for a in args), flush=True)
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} is not built")
    return subprocess.run([BIN] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                          timeout=120)


def rms_line(stdout, name):
    """The line after 'reading solutions/<name>', as the reference's test
    script reads it."""
    lines = stdout.splitlines()
    for i, line in enumerate(lines):
        if line.startswith(f"reading solutions/{name}"):
            return lines[i + 1].strip()
    raise AssertionError(f"no 'reading solutions/{name}' line in:\n{stdout}")


def last_ms(stdout):
    v = float(stdout.strip().splitlines()[-1])
    assert v >= 0.0
    return v


# reference flags (Makefile:21-71) -> get_kernel() family
FLAG_SETS = [
    ["--heuristic", "0"],                                   # base, 32x2x2, COARSEN_X 4
    ["--smem", "--heuristic", "0"],                         # smem, 32x2x2
    ["--smem", "--padded", "--heuristic", "0"],             # smem_padded
    ["--smem", "--register", "--heuristic", "1"],
    ["--coop", "--heuristic", "1"],
    ["--variant", "auto"],
    ["--coop", "--heuristic", "1", "--launch-per-pass", "--form", "3"],  # the benchmark's schedule
]


@pytest.mark.parametrize("flags", FLAG_SETS, ids=lambda f: "_".join(x.strip("-") for x in f))
def test_debug_against_oracle_solution(gpu, oracle, tmp_path, flags):
    dim, its, r = 64, 10, 1
    sol = oracle.reference_run(dim, dim, dim, 3, r, its)
    name = f"solution_{dim}_{its}_{r}_3d"
    (tmp_path / "solutions").mkdir()
    sol.tofile(tmp_path / "solutions" / name)
    res = run_cli(["--dim", dim, "--iterations", its, "--radius", r, "--debug"] + flags, tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    assert rms_line(res.stdout, name) == "rms error = 0.000000"
    last_ms(res.stdout)


@pytest.mark.parametrize("key", ["dim16_it10_r4_3d", "dim16_it6_r3_3d", "dim12_it10_r2_3d",
                                 "dim16_it10_r4_2d", "dim33_it12_r1_2d"])
@pytest.mark.parametrize("variant", ["base", "smem", "smem_padded", "smem_register", "auto"])
def test_debug_against_reference_solution_file(gpu, tmp_path, key, variant):
    """The reference's own stencil_cpu output, written as its solution file."""
    small = np.load(os.path.join(HERE, "golden", "ref_thesis_small.npz"))
    dim, it, r, d = key.replace("dim", "").replace("it", "").replace("r", "").split("_")
    dims = int(d[0])
    name = f"solution_{dim}_{it}_{r}_{dims}d"
    (tmp_path / "solutions").mkdir()
    small[key].astype(np.float32).tofile(tmp_path / "solutions" / name)
    res = run_cli(["--dim", dim, "--dimensions", dims, "--iterations", it, "--radius", r,
                   "--variant", variant, "--heuristic", "1", "--debug"], tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    assert rms_line(res.stdout, name) == "rms error = 0.000000"


def test_debug_config1_jacobi_2d(gpu, oracle, tmp_path):
    """2D 256^2 R=1 JACOBI, seeded boundary, 64 sweeps: the solution name
    carries the operator and the seed, so the THESIS solution of the same
    DIM/IT/R is never read for it."""
    dim, its, seed = 256, 64, 0x5EED0001
    u1 = oracle.init_seeded(dim, dim, 1, 2, 1, seed)
    want = oracle.run(u1, u1.copy(), 2, 1, oracle.OP_JACOBI, its).copy()
    (tmp_path / "solutions").mkdir()
    tagged = f"solution_{dim}_{its}_1_2d_jacobi_seeded{seed:x}"
    want.tofile(tmp_path / "solutions" / tagged)
    # a THESIS solution under the reference's name must not be used
    oracle.reference_run(dim, dim, 1, 2, 1, its).tofile(tmp_path / "solutions" / f"solution_{dim}_{its}_1_2d")
    res = run_cli(["--dim", dim, "--dimensions", 2, "--iterations", its, "--op", "jacobi", "--init", "seeded",
                   "--seed", hex(seed), "--variant", "auto", "--debug"], tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    assert rms_line(res.stdout, tagged) == "rms error = 0.000000"


def test_debug_wrong_solution_fails(gpu, oracle, tmp_path):
    dim, its = 48, 6
    sol = oracle.reference_run(dim, dim, dim, 3, 1, its)
    sol[20, 20, 20] += np.float32(0.01)  # sse 1e-4 > 1e-5
    name = f"solution_{dim}_{its}_1_3d"
    (tmp_path / "solutions").mkdir()
    sol.tofile(tmp_path / "solutions" / name)
    res = run_cli(["--dim", dim, "--iterations", its, "--debug"], tmp_path)
    assert res.returncode == 1
    line = rms_line(res.stdout, name)
    assert line.startswith("rms error = ") and line != "rms error = 0.000000"
    # a difference under the threshold passes (sse 1e-8 <= 1e-5)
    sol[20, 20, 20] -= np.float32(0.01)
    sol[20, 20, 20] += np.float32(1e-4)
    sol.tofile(tmp_path / "solutions" / name)
    assert run_cli(["--dim", dim, "--iterations", its, "--debug"], tmp_path).returncode == 0


def test_debug_truncated_solution_refused(gpu, oracle, tmp_path):
    dim, its = 32, 4
    (tmp_path / "solutions").mkdir()
    np.zeros(dim ** 3 - 1, np.float32).tofile(tmp_path / "solutions" / f"solution_{dim}_{its}_1_3d")
    res = run_cli(["--dim", dim, "--iterations", its, "--debug"], tmp_path)
    assert res.returncode == 1
    assert "Unable to load file: solutions/solution_32_4_1_3d" in res.stdout


def test_missing_solution_is_made_and_saved(gpu, oracle, tmp_path):
    """Without a solution file --debug makes one (BASE kernel on the GPU),
    saves it in the reference format and checks against it; --save writes
    the result file.  Both files hold the oracle's field."""
    dim, its, r = 40, 7, 2
    res = run_cli(["--dim", dim, "--iterations", its, "--radius", r, "--smem", "--padded", "--heuristic", "1",
                   "--debug", "--save"], tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    name = f"solution_{dim}_{its}_{r}_3d"
    assert f"saving solutions/{name}" in res.stdout
    assert rms_line(res.stdout, name) == "rms error = 0.000000"
    want = oracle.reference_run(dim, dim, dim, 3, r, its)
    for path in (tmp_path / "solutions" / name, tmp_path / "results" / f"result_{dim}_{its}_{r}_3d"):
        raw = path.read_bytes()
        assert len(raw) == 4 * dim ** 3
        assert np.array_equal(np.frombuffer(raw, np.uint32), want.view(np.uint32).ravel())


def test_debug_f64_and_noncubic(gpu, oracle, tmp_path):
    nx, ny, nz, its, r = 60, 44, 36, 5, 3
    u1 = oracle.init_reference(nx, ny, nz, 3, dtype=np.float64)
    want = oracle.run(u1, np.zeros_like(u1), 3, r, oracle.OP_THESIS, its).copy()
    name = f"solution_{nx}_{its}_{r}_3d_{nx}x{ny}x{nz}_f64"
    (tmp_path / "solutions").mkdir()
    want.tofile(tmp_path / "solutions" / name)
    res = run_cli(["--nx", nx, "--ny", ny, "--nz", nz, "--iterations", its, "--radius", r, "--dtype", "f64",
                   "--variant", "smem_register", "--heuristic", "1", "--debug"], tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    assert rms_line(res.stdout, name) == "rms error = 0.000000"


@pytest.mark.parametrize("args", [
    ["--smem", "--padded", "--heuristic", "0", "--radius", "2", "--block-x", "1", "--block-y", "4",
     "--block-z", "4"],                                   # padded bx < R
    ["--smem", "--padded", "--heuristic", "0", "--radius", "4", "--block-x", "8", "--block-y", "8",
     "--block-z", "2"],                                   # padded bz < R (3D)
    ["--smem", "--register", "--heuristic", "0", "--radius", "2", "--block-x", "64", "--block-y", "1"],
    ["--dim", "2"],                                       # extent not above 2R
    ["--variant", "nonsense"],
    ["--op", "laplace"],
    ["--dtype", "f16"],                                   # 16-bit storage is the mixed entry, not a sweep
    ["--no-such-flag"],
])
def test_invalid_knobs_exit_1(gpu, tmp_path, args):
    base = ["--dim", "32", "--iterations", "2"] if "--dim" not in args else ["--iterations", "2"]
    res = run_cli(base + args, tmp_path)
    assert res.returncode == 1, res.stdout + res.stderr


def test_last_line_is_kernel_ms(gpu, tmp_path):
    res = run_cli(["--dim", 128, "--iterations", 20, "--op", "jacobi", "--variant", "auto"], tmp_path)
    assert res.returncode == 0, res.stderr
    first = res.stdout.splitlines()[0]
    assert first.startswith("Version: ")
    ms = last_ms(res.stdout)
    assert 0.0 < ms < 10000.0
    assert res.stdout.strip().splitlines()[-1] == f"{ms:.4f}"
