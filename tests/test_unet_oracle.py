"""Pins for the FP64 U-Net restatement (oracle/unet_oracle.py): every
backward is checked with centred finite differences or an adjoint identity
of its forward, the forward convolution against a direct loop, and the
feature/interpolation/face functions against closed forms."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 4, 5, 3, 2))
    w = rng.standard_normal((3, 27 * 2))
    b = rng.standard_normal(3)
    y = U.conv3(x, w, b)
    N, D, H, W, Ci = x.shape
    want = np.zeros((N, D, H, W, 3))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                for co in range(3):
                    s = b[co]
                    for t, (dz, dy, dx) in enumerate(U.TAPS):
                        zz, y2, x2 = z + dz, yy + dy, xx + dx
                        if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                            s += w[co, t * Ci:(t + 1) * Ci] @ x[0, zz, y2, x2]
                    want[0, z, yy, xx, co] = s
    assert np.allclose(y, want, atol=1e-12)


def _fd(f, x, idx, h=1e-6):
    xp, xm = x.copy(), x.copy()
    xp[idx] += h
    xm[idx] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_conv3_backward_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 5, 2))
    w = rng.standard_normal((3, 54))
    b = rng.standard_normal(3)
    gy = rng.standard_normal((2, 3, 4, 5, 3))
    loss = lambda x_, w_, b_: np.sum(U.conv3(x_, w_, b_) * gy)
    gx, gw, gb = U.conv3_bwd(x, w, gy)
    for idx in [(0, 0, 0, 0, 0), (1, 2, 3, 4, 1), (0, 1, 2, 2, 0)]:
        assert abs(_fd(lambda v: loss(v, w, b), x, idx) - gx[idx]) < 1e-6
    for idx in [(0, 0), (2, 53), (1, 26)]:
        assert abs(_fd(lambda v: loss(x, v, b), w, idx) - gw[idx]) < 1e-6
    assert np.allclose(gb, gy.sum(axis=(0, 1, 2, 3)))


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 6, 2, 3))
    g = rng.standard_normal((2, 2, 3, 1, 3))
    assert abs(np.sum(U.pool(x) * g) - np.sum(x * U.pool_bwd(g))) < 1e-12
    x2 = rng.standard_normal((2, 2, 3, 1, 3))
    g2 = rng.standard_normal((2, 4, 6, 2, 3))
    assert abs(np.sum(U.up(x2) * g2) - np.sum(x2 * U.up_bwd(g2))) < 1e-12


def test_features_reproduce_linear_fields():
    n = 9
    k, j, i = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    u = 0.3 * i - 0.7 * j + 0.2 * k + 1.0
    shell_only = u.copy()
    shell_only[1:-1, 1:-1, 1:-1] = 0.0
    feat, base = U.features(shell_only[None])
    assert np.allclose(base[0], u)          # the blend of a linear field is the field
    assert np.allclose(feat[0, ..., 0], u)
    assert np.all(feat[0, 1:-1, 1:-1, 1:-1, 1] == 0)


def test_trilinear_exact_for_linear():
    m, M = 9, 17
    g = np.arange(m) / (m - 1)
    c = 2.0 * g[:, None, None] - g[None, :, None] + 0.5 * g[None, None, :]
    G = np.arange(M) / (M - 1)
    want = 2.0 * G[:, None, None] - G[None, :, None] + 0.5 * G[None, None, :]
    assert np.allclose(U.trilinear(c, M), want)


def test_fourier_faces_consistent_on_edges():
    rng = np.random.default_rng(3)
    modes = np.stack([np.stack([rng.standard_normal(8), rng.integers(1, 5, 8), rng.integers(1, 5, 8)], 1)
                      for _ in range(6)])
    u = U.fourier_field(modes, 12)
    # the sine basis vanishes on every edge of every face
    for e in (u[0, 0, :], u[0, :, 0], u[:, 0, 0], u[-1, -1, :], u[:, -1, -1]):
        assert np.max(np.abs(e)) < 1e-12
    assert np.all(u[1:-1, 1:-1, 1:-1] == 0)


def test_unet_backward_finite_differences():
    rng = np.random.default_rng(4)
    params = {}
    for name, ci, co in U.UNetRef.LAYERS:
        params[name] = (rng.standard_normal((co, 27 * ci)) * np.sqrt(2.0 / (27 * ci)),
                        rng.standard_normal(co) * 0.1)
    params["head"] = (rng.standard_normal(16) * 0.3, np.array([0.1]))
    net = U.UNetRef(params)
    x = rng.standard_normal((1, 8, 8, 8, 8))
    gy = rng.standard_normal((1, 8, 8, 8))

    def loss():
        return np.sum(net.forward(x) * gy)

    loss()
    G = net.backward(gy)
    checked = 0
    for name in ("e1a", "e2b", "btb", "d2a", "d1b"):
        w, b = params[name]
        for idx in [(0, 0), (w.shape[0] - 1, w.shape[1] - 1), (w.shape[0] // 2, 13)]:
            old = w[idx]
            w[idx] = old + 1e-6
            lp = loss()
            w[idx] = old - 1e-6
            lm = loss()
            w[idx] = old
            fd = (lp - lm) / 2e-6
            assert abs(fd - G[name][0][idx]) < 1e-5 * max(1.0, abs(fd)), (name, idx, fd, G[name][0][idx])
            checked += 1
        old = b[0]
        b[0] = old + 1e-6
        lp = loss()
        b[0] = old - 1e-6
        lm = loss()
        b[0] = old
        assert abs((lp - lm) / 2e-6 - G[name][1][0]) < 1e-5 * max(1.0, abs(G[name][1][0]))
    hw = params["head"][0]
    old = hw[3]
    hw[3] = old + 1e-6
    lp = loss()
    hw[3] = old - 1e-6
    lm = loss()
    hw[3] = old
    assert abs((lp - lm) / 2e-6 - G["head"][0][3]) < 1e-5 * max(1.0, abs(G["head"][0][3]))
    assert checked == 15


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(5).standard_normal(10000).astype(np.float32) * 37
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy().astype(np.float64)
    assert np.array_equal(U.bf16_round(x), want)
