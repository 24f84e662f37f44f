"""The band claim behind fused passes with different shells (CPU, oracle).

A fused pass of two sweeps that takes both levels' shell values from the
buffer it reads differs from the per-sweep ping-pong only at interior cells
with an index 1, 2, n-3 or n-2 on some axis (radius 1); the band fix-up
kernel k_stream_fix recomputes exactly those cells after each fused pass.
Checked here with the
oracle: the equal-shell run of the source buffer stands for the fused pass,
the ping-pong with the true shells for the reference.
"""
import numpy as np
import pytest


def band_mask(nz, ny, nx):
    def axis(n):
        idx = np.arange(n)
        return np.isin(idx, [1, 2, n - 3, n - 2]) & (idx >= 1) & (idx <= n - 2)
    k, j, i = np.meshgrid(axis(nz), axis(ny), axis(nx), indexing="ij")
    interior = np.zeros((nz, ny, nx), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    return interior & (k | j | i), interior


@pytest.mark.parametrize("shape", [(9, 11, 12), (7, 7, 8), (5, 6, 16), (3, 9, 10), (13, 4, 8)], ids=str)
@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("odd", [False, True])
def test_fused_pass_differs_only_in_band(oracle, shape, op, odd):
    nz, ny, nx = shape
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz + op)
    shell_a = oracle.init_seeded(nx, ny, nz, 3, 1, 0x5EED0050)
    shell_b = oracle.init_seeded(nx, ny, nz, 3, 1, 0x5EED0051)
    inner = rng.random((nz, ny, nx)).astype(np.float32)
    interior = np.zeros((nz, ny, nx), bool)
    interior[1:-1, 1:-1, 1:-1] = True
    # level 0: the state after an even number of sweeps (u1's shell A); an
    # odd pass reads it from u2, whose physical shell is B
    state = np.where(interior, inner, shell_a).astype(np.float32)
    src = np.where(interior, inner, shell_b if odd else shell_a).astype(np.float32)
    other = np.where(interior, 0.0, shell_b).astype(np.float32)
    want = oracle.run(state.copy(), other.copy(), 3, 1, op, 2).copy()
    fused = oracle.run(src.copy(), src.copy(), 3, 1, op, 2).copy()
    band, _ = band_mask(nz, ny, nx)
    keep = interior & ~band
    assert np.array_equal(fused[keep].view(np.uint32), want[keep].view(np.uint32))
    # the band is needed: the shells differ, so some band cell does
    if band.any():
        assert not np.array_equal(fused[band].view(np.uint32), want[band].view(np.uint32))
