"""Pins for the CPU oracle (oracle/stencil_oracle.c) before it is trusted.

Each oracle function is checked against something it does not share code
with: the reference's own outputs (golden fixtures written from the compiled
mrfarstad/thesis stencil_cpu, and oracle/_ref when it is built here), hand
values exact in binary, an independent vectorised numpy restatement, closed
forms of the discrete operator, and a dense direct solve.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_SOL = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "solutions")


def golden_records():
    with open(os.path.join(GOLD, "ref_thesis.json")) as f:
        return json.load(f)["records"]


# ------------------------------------------------------------------ golden vectors


@pytest.mark.parametrize("rec", golden_records(), ids=lambda r: r["key"])
def test_oracle_matches_reference_golden(oracle, rec):
    """Bitwise equality with the reference stencil_cpu output (SHA-256)."""
    n, d = rec["dim"], rec["dimensions"]
    u = oracle.reference_run(n, n, n, d, rec["radius"], rec["iterations"])
    raw = np.ascontiguousarray(u, np.float32).tobytes()
    assert len(raw) == rec["bytes"]
    assert hashlib.sha256(raw).hexdigest() == rec["sha256"]


def test_oracle_matches_reference_small_fields(oracle):
    small = np.load(os.path.join(GOLD, "ref_thesis_small.npz"), allow_pickle=False)
    recs = {r["key"]: r for r in golden_records()}
    assert len(small.files) >= 10
    for key in small.files:
        r = recs[key]
        n, d = r["dim"], r["dimensions"]
        u = oracle.reference_run(n, n, n, d, r["radius"], r["iterations"]).ravel()
        assert np.array_equal(u.view(np.uint32), small[key].view(np.uint32)), key


@pytest.mark.skipif(not os.path.isdir(REF_SOL), reason="oracle/_ref not built here")
def test_oracle_matches_reference_build_outputs(oracle):
    names = sorted(os.listdir(REF_SOL))
    assert names
    for name in names:
        _, dim, it, r, d = name.split("_")
        dim, it, r, dims = int(dim), int(it), int(r), int(d[0])
        ref = np.fromfile(os.path.join(REF_SOL, name), np.float32)
        mine = oracle.reference_run(dim, dim, dim, dims, r, it).ravel()
        assert np.array_equal(ref.view(np.uint32), mine.view(np.uint32)), name


# ------------------------------------------------------------------ hand values


def test_init_reference_closed_form(oracle):
    for n, d in ((512, 3), (256, 2), (7, 3)):
        if n == 512:
            n = 64  # same formula, smaller array
        u = oracle.init_reference(n, n, n, d)
        ones = int(np.sum(u == 1.0))
        assert ones == n ** d - (n - 2) ** d
        assert int(np.sum(u == 0.0)) == (n - 2) ** d


def test_thesis_known_answers_4x4(oracle):
    # 2D 4x4, R=1: interior is 2x2.  Ring 1 in u1, 0 in u2.
    # sweep 1: 2*1/4 - 0 = 0.5; sweep 2: (2*0 + 2*0.5)/4 - 0.5 = -0.25;
    # sweep 3: (2*1 + 2*-0.25)/4 + 0.25 = 0.625  (all exact in binary)
    for it, val in ((1, 0.5), (2, -0.25), (3, 0.625)):
        u = oracle.reference_run(4, 4, 1, 2, 1, it)
        assert np.all(u[1:3, 1:3] == np.float32(val)), (it, u)


def test_seeded_generator_values(oracle):
    # splitmix64 of the published algorithm, top 24 bits
    def splitmix(z):
        m = (1 << 64) - 1
        z = (z + 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)

    m = (1 << 64) - 1
    for seed in (0, 0x5EED0002, 12345):
        for idx in (0, 1, 7, 1 << 33, 134217727):
            h = splitmix(seed ^ ((idx * 0x9E3779B97F4A7C15) & m))
            want = (h >> 40) / 2.0 ** 24
            assert oracle.uniform(seed, idx) == np.float32(want)
    u = oracle.init_seeded(10, 9, 8, 3, 2, 0x5EED0002)
    assert np.all(u[2:-2, 2:-2, 2:-2] == 0)
    shell = np.ones(u.shape, bool)
    shell[2:-2, 2:-2, 2:-2] = False
    assert np.all((u[shell] >= 0) & (u[shell] < 1)) and np.count_nonzero(u[shell]) > 0.95 * shell.sum()
    idx = 3 + 9 * 10 * 1 + 10 * 0  # (i=3, j=0, k=1)
    assert u.ravel()[idx] == oracle.uniform(0x5EED0002, idx)


# ------------------------------------------------------------------ independent numpy restatement


def numpy_sweep(u, dims, r, op, w):
    """Vectorised float32 restatement: same per-cell operation order."""
    out = u.copy()
    f = np.float32
    if dims == 3:
        nz, ny, nx = u.shape
        c = u[r:nz - r, r:ny - r, r:nx - r]
        acc = np.zeros_like(c)
        sl = lambda dk, dj, di: u[r + dk:nz - r + dk, r + dj:ny - r + dj, r + di:nx - r + di]
        order = [(0, 0, -d) for d in range(r, 0, -1)] + [(0, 0, d) for d in range(1, r + 1)] \
            + [(0, -d, 0) for d in range(r, 0, -1)] + [(0, d, 0) for d in range(1, r + 1)] \
            + [(-d, 0, 0) for d in range(r, 0, -1)] + [(d, 0, 0) for d in range(1, r + 1)]
        for o in order:
            acc = (acc + sl(*o)).astype(np.float32)
        res = (acc / f(2 * dims * r) - c) if op == 0 else acc * f(w)
        out[r:nz - r, r:ny - r, r:nx - r] = res
    else:
        ny, nx = u.shape
        c = u[r:ny - r, r:nx - r]
        acc = np.zeros_like(c)
        sl = lambda dj, di: u[r + dj:ny - r + dj, r + di:nx - r + di]
        order = [(0, -d) for d in range(r, 0, -1)] + [(0, d) for d in range(1, r + 1)] \
            + [(-d, 0) for d in range(r, 0, -1)] + [(d, 0) for d in range(1, r + 1)]
        for o in order:
            acc = (acc + sl(*o)).astype(np.float32)
        res = (acc / f(2 * dims * r) - c) if op == 0 else acc * f(w)
        out[r:ny - r, r:nx - r] = res
    return out


@pytest.mark.parametrize("dims,shape,r,op,its", [
    (2, (13, 17), 1, 0, 5), (2, (19, 23), 3, 0, 4), (2, (16, 21), 2, 1, 9),
    (3, (9, 11, 13), 1, 0, 4), (3, (12, 10, 11), 2, 1, 6), (3, (14, 13, 12), 4, 1, 3),
    (3, (10, 9, 12), 3, 0, 3),
])
def test_oracle_vs_numpy_restatement(oracle, dims, shape, r, op, its):
    if dims == 3:
        nz, ny, nx = shape
    else:
        (ny, nx), nz = shape, 1
    u1 = oracle.init_seeded(nx, ny, nz, dims, r, 0xABC)
    rng = np.random.default_rng(5)
    u1 = u1 + (rng.random(u1.shape, dtype=np.float32) * (u1 == 0))  # random interior too
    u1 = u1.astype(np.float32)
    u2 = u1.copy()
    w = oracle.default_weight(dims, r)
    a, b = u1.copy(), u2.copy()
    for _ in range(its):
        b = numpy_sweep(a, dims, r, op, w)
        a, b = b, a
    got = oracle.run(u1, u2, dims, r, op, its)
    assert np.array_equal(got.view(np.uint32), a.view(np.uint32))


def test_oracle_thread_count_bitwise(oracle):
    u1 = oracle.init_seeded(40, 33, 37, 3, 2, 7)
    v1 = u1.copy()
    a = oracle.run(u1, u1.copy(), 3, 2, 1, 7, nthreads=1)
    b = oracle.run(v1, v1.copy(), 3, 2, 1, 7, nthreads=8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ------------------------------------------------------------------ JACOBI closed forms


def test_jacobi_constant_field_fixed_point(oracle):
    for dims, r in ((2, 1), (3, 1), (3, 4), (2, 2)):
        shape = (12, 13, 14) if dims == 3 else (15, 16)
        u = np.full(shape, 0.75, np.float32)  # 0.75 * 2DR * (1/2DR): exact for 2D R=1
        out = oracle.run(u.copy(), u.copy(), dims, r, 1, 3)
        ulp = np.spacing(np.float32(0.75))
        assert np.max(np.abs(out - 0.75)) <= 4 * ulp
        ud = u.astype(np.float64)
        outd = oracle.run(ud.copy(), ud.copy(), dims, r, 1, 3)
        assert np.max(np.abs(outd - 0.75)) < 1e-15


def test_jacobi_linear_field_preserved(oracle):
    k, j, i = np.meshgrid(np.arange(10), np.arange(11), np.arange(12), indexing="ij")
    u = (0.3 * i - 0.2 * j + 0.1 * k + 0.5).astype(np.float64)
    out = oracle.run(u.copy(), u.copy(), 3, 1, 1, 5)
    assert np.max(np.abs(out - u)) < 1e-14
    out2 = oracle.run(u.copy(), u.copy(), 3, 2, 1, 5)  # wider star, still exact for linear
    assert np.max(np.abs(out2 - u)) < 1e-14


def test_jacobi_discrete_harmonic_fixed_point(oracle):
    # u = sin(a i) sin(b j) sinh(c k), cosh c = 3 - cos a - cos b  is an exact
    # zero of the 7-point Laplacian: one JACOBI sweep leaves it unchanged.
    a, b = 0.3, 0.45
    c = np.arccosh(3 - np.cos(a) - np.cos(b))
    k, j, i = np.meshgrid(np.arange(14), np.arange(12), np.arange(10), indexing="ij")
    u = np.sin(a * i) * np.sin(b * j) * np.sinh(c * k)
    out = oracle.run(u.copy(), u.copy(), 3, 1, 1, 1)
    assert np.max(np.abs(out - u)) < 1e-12 * np.max(np.abs(u))
    assert oracle.residual_norm(u, 3, 1) < 1e-11 * np.linalg.norm(u)
    # a wrong operator fails this check: a mis-signed weight does not fix u
    bad = u.copy()
    oracle.sweep(u, bad, 3, 1, 1, weight=-1.0 / 6.0)
    assert np.max(np.abs(bad - u)) > 1e-2


def _dense_dirichlet_solution(u, dims):
    """Solve the discrete Laplace Dirichlet problem with a dense solver."""
    if dims == 2:
        ny, nx = u.shape
        idx = -np.ones(u.shape, int)
        cells = [(j, i) for j in range(1, ny - 1) for i in range(1, nx - 1)]
        for n, (j, i) in enumerate(cells):
            idx[j, i] = n
        A = np.zeros((len(cells), len(cells)))
        b = np.zeros(len(cells))
        for n, (j, i) in enumerate(cells):
            A[n, n] = 4
            for dj, di in ((0, 1), (0, -1), (1, 0), (-1, 0)):
                jj, ii = j + dj, i + di
                if idx[jj, ii] >= 0:
                    A[n, idx[jj, ii]] -= 1
                else:
                    b[n] += u[jj, ii]
        x = np.linalg.solve(A, b)
        sol = u.copy()
        for n, (j, i) in enumerate(cells):
            sol[j, i] = x[n]
        return sol
    raise NotImplementedError


def test_jacobi_converges_to_direct_solve(oracle):
    u = oracle.init_seeded(14, 12, 1, 2, 1, 99, dtype=np.float64)
    sol = _dense_dirichlet_solution(u, 2)
    out = oracle.run(u.copy(), u.copy(), 2, 1, 1, 2000)
    assert np.max(np.abs(out - sol)) < 1e-10
    # maximum principle along the way
    mid = oracle.run(u.copy(), u.copy(), 2, 1, 1, 37)
    assert mid.min() >= 0 and mid.max() <= u.max()


def test_residual_norm_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    u = rng.random((9, 10, 11))
    r = 1
    s = (u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1]
         + u[:-2, 1:-1, 1:-1] + u[2:, 1:-1, 1:-1])
    want = np.sqrt(np.sum((s / 6.0 - u[1:-1, 1:-1, 1:-1]) ** 2))
    assert abs(oracle.residual_norm(u, 3, r) - want) < 1e-12 * want
    uf = u.astype(np.float32)
    want32 = np.sqrt(np.sum(((uf[1:-1, 1:-1, :-2].astype(np.float64) + uf[1:-1, 1:-1, 2:]
                               + uf[1:-1, :-2, 1:-1] + uf[1:-1, 2:, 1:-1] + uf[:-2, 1:-1, 1:-1]
                               + uf[2:, 1:-1, 1:-1]) / 6.0 - uf[1:-1, 1:-1, 1:-1]) ** 2))
    assert abs(oracle.residual_norm(uf, 3, r) - want32) < 1e-12 * want32


def test_sse_closed_form(oracle):
    a = np.zeros(1000, np.float32)
    b = a.copy()
    b[::10] = 0.5  # 100 cells differ by 0.5
    assert oracle.sse(a, b) == 25.0
    assert oracle.sse(a, a) == 0.0
    # the square is rounded to FP32 before the FP64 sum
    # (include/stencil_error_checker.h:15): (1 + 2^-12)^2 = 1 + 2^-11 + 2^-24
    # is a tie in FP32 and rounds to even, 1 + 2^-11; squared in FP64 it
    # would keep the 2^-24 term
    d = np.float32(1 + 2.0 ** -12)
    assert oracle.sse(np.array([d], np.float32), np.zeros(1, np.float32)) == 1 + 2.0 ** -11
    assert (float(d) * float(d)) != 1 + 2.0 ** -11


# ------------------------------------------------------------------ 16-bit storage


def test_f16_rounding_matches_numpy(oracle):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504, 65520, 65519.99, 1e-8, 3e-8, 6e-8, 5.96e-8, 2.98e-8,
                  6.1e-5, np.inf, -np.inf], np.float32)])
    x = x.astype(np.float32)
    h = oracle.round16(x, True)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(h, want)
    back = oracle.widen16(h, True)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))


def test_bf16_rounding_matches_torch(oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(12)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    x = np.concatenate([x, np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.4e38], np.float32)])
    h = oracle.round16(x, False)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(h, want)
    back = oracle.widen16(h, False)
    assert np.array_equal(back, torch.from_numpy(want.view(np.int16)).view(torch.bfloat16).float().numpy())


def numpy_mixed(x0, dims, r, its, restart, is_f16, w):
    """Independent restatement of the mixed-precision scheme with numpy's
    float16 or torch's bfloat16 rounding."""
    if is_f16:
        rnd = lambda a: a.astype(np.float32).astype(np.float16).astype(np.float32)
    else:
        import torch
        rnd = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).float().numpy()
    x = x0.astype(np.float32).copy()
    interior = np.zeros(x.shape, bool)
    interior[(slice(r, -r),) * dims] = True

    def G(u):  # FP32 JACOBI cell on the interior (0 elsewhere)
        return np.where(interior, numpy_sweep(u, dims, r, 1, w), 0).astype(np.float32)

    d = np.where(interior, rnd(G(x) - x), 0).astype(np.float32)
    e = np.zeros_like(x)
    for it in range(1, its + 1):
        z = np.zeros_like(x)
        s = numpy_sweep(e, dims, r, 1, w)  # s*w on interior
        e = np.where(interior, rnd((d + s).astype(np.float32)), 0).astype(np.float32)
        if it % restart == 0 and it < its:
            x = (x + e).astype(np.float32)
            e = z
            d = np.where(interior, rnd(G(x) - x), 0).astype(np.float32)
    return (x + e).astype(np.float32)


@pytest.mark.parametrize("is_f16", [False, True])
def test_mixed_oracle_vs_numpy(oracle, is_f16):
    pytest.importorskip("torch")
    x0 = oracle.init_seeded(12, 11, 10, 3, 1, 0x77)
    w = oracle.default_weight(3, 1)
    got, folds = oracle.mixed_run(x0, 3, 1, 23, 5, is_f16)
    want = numpy_mixed(x0, 3, 1, 23, 5, is_f16, w)
    assert folds == 4
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_mixed_error_vs_fp64(oracle):
    x0 = oracle.init_seeded(24, 22, 20, 3, 1, 0x99)
    ref = oracle.run(x0.astype(np.float64), x0.astype(np.float64), 3, 1, 1, 60)
    f32 = oracle.run(x0.copy(), x0.copy(), 3, 1, 1, 60)
    bf, _ = oracle.mixed_run(x0, 3, 1, 60, 10, False)
    hf, _ = oracle.mixed_run(x0, 3, 1, 60, 10, True)
    e32 = np.max(np.abs(f32 - ref))
    ebf = np.max(np.abs(bf - ref))
    ehf = np.max(np.abs(hf - ref))
    assert e32 < 1e-6
    assert ebf < 2e-3 and ehf < 3e-4
    # without the fold the 16-bit error is far larger: plain BF16 storage
    torch = pytest.importorskip("torch")
    u = torch.from_numpy(x0).to(torch.bfloat16).float().numpy()
    for _ in range(60):
        u = torch.from_numpy(numpy_sweep(u, 3, 1, 1, oracle.default_weight(3, 1))).to(torch.bfloat16).float().numpy()
    assert np.max(np.abs(u - ref)) > 2 * ebf
