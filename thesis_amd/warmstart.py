"""Learned warm start for 3D Laplace Jacobi solves.

A 3-level 3D convolutional U-Net (16/32/64 channels, 3x3x3 convolutions,
BF16 weights and activations, FP32 accumulation) maps the Dirichlet data of
a 64^3 problem to a correction of the transfinite blend of its faces.  It is
trained on the GPU, with the kernels of thesis_amd/csrc/unet.hip, on
converged solves that the library's own batched Jacobi sweep produces.  At
solve time the prediction is trilinearly interpolated to the target grid
(128^3) and plain JACOBI sweeps take it to the requested residual.

Everything here is host orchestration; all arithmetic runs in the HIP
library.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import (BATCH_PERSISTENT, DeviceArray, INIT_SEEDED, OP_JACOBI, StencilParams, Stream,
               dispatch_batched, dispatch_kernels, residual_norm, synchronize)
from . import nn, workloads

MODES = 8
FEAT_C = 8
LAYERS = [("e1a", 8, 16), ("e1b", 16, 16), ("e2a", 16, 32), ("e2b", 32, 32),
          ("bta", 32, 64), ("btb", 64, 64), ("d2a", 96, 32), ("d2b", 32, 32),
          ("d1a", 48, 16), ("d1b", 16, 16)]


def face_modes(count: int, seed: int) -> np.ndarray:
    """(count, 6, MODES, 3) float32 array of (a, kx, ky) per face: problem i
    uses workloads.fourier_faces(seed + i)."""
    out = np.zeros((count, 6, MODES, 3), np.float32)
    for i in range(count):
        f = workloads.fourier_faces(0, seed + i, modes=MODES)
        for fi, name in enumerate(("x0", "x1", "y0", "y1", "z0", "z1")):
            out[i, fi] = np.array(f[name], np.float32)
    return out


class UNet:
    """The network with its parameters, Adam state and per-batch buffers."""

    def __init__(self, batch: int, n: int = 64, seed: int = 1, stream=None):
        assert n % 4 == 0
        self.B, self.n = batch, n
        self.s = stream or Stream()
        rng = np.random.default_rng(seed)
        self.params = {}
        for name, ci, co in LAYERS:
            kp = nn.kpad(ci)
            w = np.zeros((co, kp), np.float32)
            w[:, :27 * ci] = rng.standard_normal((co, 27 * ci)) * math.sqrt(2.0 / (27 * ci))
            self._add_conv(name, ci, co, w, np.zeros(co, np.float32))
        hw = (rng.standard_normal(16) * 1e-3).astype(np.float32)
        self._add_head(hw, np.zeros(1, np.float32))
        self.step_count = 0
        self._alloc()

    # -- parameters
    def _add_conv(self, name, ci, co, w, b):
        kp = nn.kpad(ci)
        d = dict(ci=ci, co=co, kpad=kp, kpadT=nn.kpad(co),
                 w=DeviceArray.from_numpy(w), b=DeviceArray.from_numpy(b),
                 wbf=DeviceArray((co, kp), np.uint16), wt=DeviceArray((ci, nn.kpad(co)), np.uint16),
                 gw=DeviceArray((co, kp), np.float32), gb=DeviceArray((co,), np.float32),
                 mw=DeviceArray((co, kp), np.float32), vw=DeviceArray((co, kp), np.float32),
                 mb=DeviceArray((co,), np.float32), vb=DeviceArray((co,), np.float32))
        for k in ("mw", "vw", "mb", "vb"):
            d[k].fill_zero()
        self.params[name] = d
        self._refresh(name)

    def _add_head(self, w, b):
        d = dict(w=DeviceArray.from_numpy(w), b=DeviceArray.from_numpy(b),
                 gw=DeviceArray((16,), np.float32), gb=DeviceArray((1,), np.float32),
                 mw=DeviceArray((16,), np.float32), vw=DeviceArray((16,), np.float32),
                 mb=DeviceArray((1,), np.float32), vb=DeviceArray((1,), np.float32))
        for k in ("mw", "vw", "mb", "vb"):
            d[k].fill_zero()
        self.params["head"] = d

    def _refresh(self, name):
        d = self.params[name]
        nn.f32_to_bf16(d["w"], d["wbf"], d["co"] * d["kpad"], self.s)
        nn.wt_transpose(d["wbf"], d["wt"], d["ci"], d["co"], self.s)

    def state_dict(self):
        out = {}
        for name, d in self.params.items():
            if name == "head":
                out["head.w"] = d["w"].numpy(self.s)
                out["head.b"] = d["b"].numpy(self.s)
            else:
                out[name + ".w"] = d["w"].numpy(self.s)[:, :27 * d["ci"]]
                out[name + ".b"] = d["b"].numpy(self.s)
        return out

    def load_state_dict(self, sd):
        for name, d in self.params.items():
            if name == "head":
                d["w"].upload(np.asarray(sd["head.w"], np.float32), self.s)
                d["b"].upload(np.asarray(sd["head.b"], np.float32), self.s)
            else:
                w = np.zeros((d["co"], d["kpad"]), np.float32)
                w[:, :27 * d["ci"]] = sd[name + ".w"]
                d["w"].upload(w, self.s)
                d["b"].upload(np.asarray(sd[name + ".b"], np.float32), self.s)
                self._refresh(name)

    def save(self, path):
        np.savez(path, **self.state_dict())

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.load_state_dict({k: z[k] for k in z.files})

    # -- buffers
    def _alloc(self):
        B, n = self.B, self.n
        self.sh1, self.sh2, self.sh3 = (B, n, n, n), (B, n // 2, n // 2, n // 2), (B, n // 4, n // 4, n // 4)
        self.P1, self.P2, self.P3 = B * n ** 3, B * (n // 2) ** 3, B * (n // 4) ** 3
        bf = lambda P, C: DeviceArray((P, C), np.uint16)
        f32 = lambda *s: DeviceArray(s, np.float32)
        P1, P2, P3 = self.P1, self.P2, self.P3
        self.a = dict(x0=bf(P1, 8), e1a=bf(P1, 16), e1b=bf(P1, 16), p1=bf(P2, 16), e2a=bf(P2, 32),
                      e2b=bf(P2, 32), p2=bf(P3, 32), bta=bf(P3, 64), btb=bf(P3, 64), cat2=bf(P2, 96),
                      d2a=bf(P2, 32), d2b=bf(P2, 32), cat1=bf(P1, 48), d1a=bf(P1, 16), d1b=bf(P1, 16))
        self.y, self.base, self.pred, self.gu = f32(P1), f32(P1), f32(P1), f32(P1)
        self.u = f32(B, n, n, n)
        # FP32 gradient buffers per level: A = input-gradient of the level's
        # convolutions (widest: the concatenations), B = output-gradients
        self.gA1, self.gB1, self.ge1b = f32(P1, 48), f32(P1, 16), f32(P1, 16)
        self.gA2, self.gB2, self.ge2b = f32(P2, 96), f32(P2, 32), f32(P2, 32)
        self.gA3, self.gB3 = f32(P3, 64), f32(P3, 64)
        self.gb1, self.gb2, self.gb3 = bf(P1, 16), bf(P2, 32), bf(P3, 64)
        self.loss = DeviceArray((1,), np.float64)
        self.sloss = DeviceArray((1,), np.float64)
        self.e1, self.e2, self.ge = f32(B, n, n, n), f32(B, n, n, n), f32(P1)

    def load_fields(self, src: DeviceArray, indices):
        """Copy fields src[i] (n^3 each) for i in indices into the batch."""
        nb = self.n ** 3 * 4
        for b, i in enumerate(indices):
            _dtod(self.u.ptr.value + b * nb, src.ptr.value + int(i) * nb, nb, self.s)

    # -- forward
    def _conv(self, name, x, out, shape, relu=True):
        d = self.params[name]
        nn.conv3(x, d["wbf"], d["b"], out, None, shape, d["ci"], d["co"], relu=relu, stream=self.s)

    def forward(self, target=None):
        """Prediction (blend + network correction on the interior, boundary on
        the shell) of the fields in self.u into self.pred.  With a target the
        MSE over the interior and its gradient are formed (self.loss)."""
        a, s = self.a, self.s
        nn.features(self.u, a["x0"], self.base, self.B, self.n, FEAT_C, s)
        self._conv("e1a", a["x0"], a["e1a"], self.sh1)
        self._conv("e1b", a["e1a"], a["e1b"], self.sh1)
        nn.pool_fwd(a["e1b"], a["p1"], self.sh2, 16, s)
        self._conv("e2a", a["p1"], a["e2a"], self.sh2)
        self._conv("e2b", a["e2a"], a["e2b"], self.sh2)
        nn.pool_fwd(a["e2b"], a["p2"], self.sh3, 32, s)
        self._conv("bta", a["p2"], a["bta"], self.sh3)
        self._conv("btb", a["bta"], a["btb"], self.sh3)
        nn.up_fwd(a["btb"], a["cat2"], self.sh3, 64, 96, 0, s)
        nn.chan_copy(a["e2b"], a["cat2"], self.P2, 32, 32, 0, 96, 64, False, stream=s)
        self._conv("d2a", a["cat2"], a["d2a"], self.sh2)
        self._conv("d2b", a["d2a"], a["d2b"], self.sh2)
        nn.up_fwd(a["d2b"], a["cat1"], self.sh2, 32, 48, 0, s)
        nn.chan_copy(a["e1b"], a["cat1"], self.P1, 16, 16, 0, 48, 32, False, stream=s)
        self._conv("d1a", a["cat1"], a["d1a"], self.sh1)
        self._conv("d1b", a["d1a"], a["d1b"], self.sh1)
        h = self.params["head"]
        nn.head_fwd(a["d1b"], h["w"], h["b"], self.y, self.P1, 16, s)
        if target is not None:
            self.loss.fill_zero(s)
            count = self.B * (self.n - 2) ** 3
            nn.mse(self.base, self.y, target, self.gu, self.loss, self.pred, self.B, self.n,
                   1.0 / count, s)
        else:
            nn.mse(self.base, self.y, None, None, None, self.pred, self.B, self.n, 1.0, s)

    def loss_value(self) -> float:
        return float(self.loss.numpy(self.s)[0]) / (self.B * (self.n - 2) ** 3)

    def smooth_loss_value(self) -> float:
        return float(self.sloss.numpy(self.s)[0]) / (self.B * (self.n - 2) ** 3)

    def _smooth(self, x: DeviceArray, sweeps: int):
        """x <- G0^sweeps x, G0 the JACOBI sweep with a zero shell (x's shell is
        zero), on all B samples with the batched persistent kernel."""
        nb = self.n ** 3 * 4
        self.e2.copy_from(x, self.s)
        probs = [dict(nx=self.n, ny=self.n, nz=self.n, radius=1, iterations=sweeps, op=OP_JACOBI,
                      u1=x.ptr.value + b * nb, u2=self.e2.ptr.value + b * nb) for b in range(self.B)]
        dispatch_batched(probs, [], BATCH_PERSISTENT, stream=self.s)

    def add_smoothed_error_loss(self, target: DeviceArray, sweeps: int, weight: float):
        """Adds weight * ||G0^K (pred - target)||^2 / count to the objective
        (its gradient G0^K applied to 2 G0^K e, G0 being symmetric) on top of
        the MSE gradient already in self.gu.  The error after K zero-boundary
        JACOBI sweeps keeps the slowly decaying modes that set the number of
        sweeps a warm start needs."""
        P, s = self.P1, self.s
        count = self.B * (self.n - 2) ** 3
        nn.axpby(self.e1, 1.0, self.pred, -1.0, target, P, s)
        self._smooth(self.e1, sweeps)
        self.sloss.fill_zero(s)
        nn.sumsq(self.e1, self.ge, self.sloss, P, weight / count, s)
        self._smooth(self.ge, sweeps)
        nn.axpby(self.gu, 1.0, self.gu, 1.0, self.ge, P, s)

    # -- backward
    def _layer_bwd(self, name, x_in, y_out, g_out, gb16, shape, P, g_in):
        """Given the FP32 gradient of the layer's (post-ReLU) output, add the
        layer's weight and bias gradients and write the FP32 gradient of its
        input into g_in (skipped when g_in is None)."""
        d, s = self.params[name], self.s
        nn.relu_bwd(g_out, y_out, gb16, P * d["co"], s)
        nn.conv3_wgrad(x_in, gb16, d["gw"], shape, d["ci"], d["co"], s)
        nn.bias_grad(gb16, d["gb"], P, d["co"], s)
        if g_in is not None:
            nn.conv3(gb16, d["wt"], None, None, g_in, shape, d["co"], d["ci"], stream=s)

    def backward(self):
        a, s = self.a, self.s
        for d in self.params.values():
            d["gw"].fill_zero(s)
            d["gb"].fill_zero(s)
        h = self.params["head"]
        nn.head_bwd(a["d1b"], h["w"], self.gu, self.gB1, h["gw"], h["gb"], self.P1, 16, s)
        # decoder level 1
        self._layer_bwd("d1b", a["d1a"], a["d1b"], self.gB1, self.gb1, self.sh1, self.P1, self.gA1)
        self._layer_bwd("d1a", a["cat1"], a["d1a"], self.gA1, self.gb1, self.sh1, self.P1, self.gA1)
        nn.chan_copy(self.gA1, self.ge1b, self.P1, 16, 48, 32, 16, 0, True, stream=s)
        nn.up_bwd(self.gA1, self.gB2, self.sh2, 32, 48, 0, s)
        # decoder level 2
        self._layer_bwd("d2b", a["d2a"], a["d2b"], self.gB2, self.gb2, self.sh2, self.P2, self.gA2)
        self._layer_bwd("d2a", a["cat2"], a["d2a"], self.gA2, self.gb2, self.sh2, self.P2, self.gA2)
        nn.chan_copy(self.gA2, self.ge2b, self.P2, 32, 96, 64, 32, 0, True, stream=s)
        nn.up_bwd(self.gA2, self.gB3, self.sh3, 64, 96, 0, s)
        # bottleneck
        self._layer_bwd("btb", a["bta"], a["btb"], self.gB3, self.gb3, self.sh3, self.P3, self.gA3)
        self._layer_bwd("bta", a["p2"], a["bta"], self.gA3, self.gb3, self.sh3, self.P3, self.gA3)
        nn.pool_bwd(self.gA3, self.ge2b, self.sh3, 32, s)
        # encoder level 2
        self._layer_bwd("e2b", a["e2a"], a["e2b"], self.ge2b, self.gb2, self.sh2, self.P2, self.gA2)
        self._layer_bwd("e2a", a["p1"], a["e2a"], self.gA2, self.gb2, self.sh2, self.P2, self.gA2)
        nn.pool_bwd(self.gA2, self.ge1b, self.sh2, 16, s)
        # encoder level 1
        self._layer_bwd("e1b", a["e1a"], a["e1b"], self.ge1b, self.gb1, self.sh1, self.P1, self.gA1)
        self._layer_bwd("e1a", a["x0"], a["e1a"], self.gA1, self.gb1, self.sh1, self.P1, None)

    def grads(self):
        """Host copies of the parameter gradients (name -> (gw, gb))."""
        out = {}
        for name, d in self.params.items():
            gw = d["gw"].numpy(self.s)
            if name != "head":
                gw = gw[:, :27 * d["ci"]]
            out[name] = (gw, d["gb"].numpy(self.s))
        return out

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        s = self.s
        for name, d in self.params.items():
            nw = int(np.prod(d["w"].shape))
            nn.adam(d["w"], d["gw"], d["mw"], d["vw"], d.get("wbf"), nw, lr, b1, b2, eps,
                    self.step_count, s)
            nn.adam(d["b"], d["gb"], d["mb"], d["vb"], None, int(np.prod(d["b"].shape)), lr, b1, b2,
                    eps, self.step_count, s)
            if name != "head":
                nn.wt_transpose(d["wbf"], d["wt"], d["ci"], d["co"], s)


def _dtod(dst: int, src: int, nbytes: int, stream):
    import ctypes as C

    from . import _check, _sh, lib
    _check(lib().st_memcpy_dtod(C.c_void_p(dst), C.c_void_p(src), nbytes, _sh(stream)))


# ------------------------------------------------------------------ data

def generate_solves(count: int, n: int, seed: int, sweeps: int, stream=None):
    """`count` converged n^3 Laplace solves: faces from face_modes(count, seed),
    interior started from the transfinite blend of the faces, then `sweeps`
    JACOBI sweeps of the batched persistent kernel (an even count, so the
    result is in the first buffer).  Returns (u, modes) with u a DeviceArray
    (count, n, n, n)."""
    assert sweeps % 2 == 0
    s = stream or Stream()
    modes = face_modes(count, seed)
    dm = DeviceArray.from_numpy(modes, s)
    u1 = DeviceArray((count, n, n, n), np.float32)
    u2 = DeviceArray((count, n, n, n), np.float32)
    u1.fill_zero(s)
    nn.fourier_faces(u1, dm, count, n, MODES, s)
    chunk = 64
    feat = DeviceArray((chunk * n ** 3, FEAT_C), np.uint16)
    base = DeviceArray((chunk, n, n, n), np.float32)
    nb = n ** 3 * 4
    for c0 in range(0, count, chunk):
        c = min(chunk, count - c0)
        import ctypes as C
        nn.features(C.c_void_p(u1.ptr.value + c0 * nb), feat, base, c, n, FEAT_C, s)
        _dtod(u1.ptr.value + c0 * nb, base.ptr.value, c * nb, s)
    feat.free()
    base.free()
    _dtod(u2.ptr.value, u1.ptr.value, count * nb, s)
    probs = [dict(nx=n, ny=n, nz=n, radius=1, iterations=sweeps, op=OP_JACOBI,
                  u1=u1.ptr.value + i * nb, u2=u2.ptr.value + i * nb) for i in range(count)]
    dispatch_batched(probs, [], BATCH_PERSISTENT, stream=s)
    s.sync()
    u2.free()
    dm.free()
    return u1, modes


def relative_residual(u: DeviceArray, n: int, r0: float, stream=None) -> float:
    p = StencilParams.make(dim=n, radius=1, op=OP_JACOBI)
    return residual_norm(p, u, stream) / r0


# ------------------------------------------------------------------ training

def train(net: UNet, fields: DeviceArray, count: int, steps: int, lr: float = 1e-3,
          lr_final: float = 1e-4, seed: int = 0, log_every: int = 50, verbose: bool = False,
          smooth_sweeps: int = 0, smooth_weight: float = 0.0):
    """Adam on random batches of `fields` (the converged solves); the learning
    rate decays geometrically from lr to lr_final.  With smooth_sweeps > 0 the
    objective adds smooth_weight * ||G0^K (pred - target)||^2.  Returns
    [(step, mse)] or [(step, mse, smoothed)]."""
    rng = np.random.default_rng(seed)
    order, pos, log = rng.permutation(count), 0, []
    t0 = time.perf_counter()
    for step in range(steps):
        if pos + net.B > count:
            order, pos = rng.permutation(count), 0
        idx = order[pos:pos + net.B]
        pos += net.B
        net.load_fields(fields, idx)
        net.forward(target=net.u)
        if smooth_sweeps:
            net.add_smoothed_error_loss(net.u, smooth_sweeps, smooth_weight)
        net.backward()
        rate = lr * (lr_final / lr) ** (step / max(1, steps - 1))
        net.adam_step(rate)
        if step % log_every == 0 or step == steps - 1:
            loss = net.loss_value()
            rec = (step, loss) if not smooth_sweeps else (step, loss, net.smooth_loss_value())
            log.append(rec)
            if verbose:
                print(f"[train] step {step} loss {rec[1:]} lr {rate:.2e} "
                      f"{time.perf_counter() - t0:.1f}s", flush=True)
    return log


# ------------------------------------------------------------------ solving

def jacobi_to_tolerance(u1: DeviceArray, u2: DeviceArray, n: int, r0: float, tol: float,
                        check_every: int, max_sweeps: int, stream):
    """JACOBI sweeps (the library's fastest single-grid kernel) on u1/u2 until
    ||r|| / r0 <= tol, checking every `check_every` sweeps (even, so the field
    stays in u1).  Returns (sweeps, relative residual)."""
    assert check_every % 2 == 0
    p = StencilParams.make(dim=n, radius=1, iterations=check_every, op=OP_JACOBI, variant="auto",
                           heuristic=1)
    q = StencilParams.make(dim=n, radius=1, op=OP_JACOBI)
    sweeps, rel = 0, residual_norm(q, u1, stream) / r0
    while rel > tol and sweeps < max_sweeps:
        dispatch_kernels(p, u1, u2, stream)
        sweeps += check_every
        rel = residual_norm(q, u1, stream) / r0
    return sweeps, rel


def coarse_refine(u: DeviceArray, scratch: DeviceArray, m: int, sweeps: int, stream):
    """`sweeps` (even) JACOBI sweeps of the m^3 field u in place (scratch is
    the second buffer; both hold u's shell)."""
    if sweeps <= 0:
        return
    assert sweeps % 2 == 0
    scratch.copy_from(u, stream)
    p = StencilParams.make(dim=m, radius=1, iterations=sweeps, op=OP_JACOBI, variant="auto", heuristic=1)
    in_u2, _ = dispatch_kernels(p, u, scratch, stream)
    assert not in_u2


def compare_warm_start(state: dict, seed: int, n: int = 128, m: int = 64, tol: float = 1e-6,
                       check_every: int = 200, max_sweeps: int = 200000, stream=None,
                       coarse_sweeps: int = 0):
    """Plain zero-start JACOBI versus the learned warm start on one 128^3
    problem with Fourier-mode faces (seed).  Relative residuals use the
    zero-start residual norm as the common denominator.  With coarse_sweeps
    > 0, two more starts are timed: the network's m^3 prediction refined by
    that many m^3 JACOBI sweeps before the upsampling ("learned_coarse"),
    and the same refinement of the network's input guess alone
    ("blend_coarse", the control without the network).  Returns a dict."""
    s = stream or Stream()
    modes = DeviceArray.from_numpy(face_modes(1, seed), s)
    u1 = DeviceArray((n, n, n), np.float32)
    u2 = DeviceArray((n, n, n), np.float32)
    u1.fill_zero(s)
    nn.fourier_faces(u1, modes, 1, n, MODES, s)
    u2.copy_from(u1, s)
    q = StencilParams.make(dim=n, radius=1, op=OP_JACOBI)
    r0 = residual_norm(q, u1, s)
    out = {"n": n, "coarse": m, "tol": tol, "check_every": check_every, "seed": seed, "r0": r0}
    # plain zero start
    s.sync()
    t0 = time.perf_counter()
    sw, rel = jacobi_to_tolerance(u1, u2, n, r0, tol, check_every, max_sweeps, s)
    s.sync()
    out["plain"] = {"sweeps": sw, "seconds": time.perf_counter() - t0, "rel_residual": rel}
    # learned: network at m^3, trilinear to n^3, then the same sweeps
    net = UNet(1, m, stream=s)
    net.load_state_dict(state)
    net.u.fill_zero(s)
    nn.fourier_faces(net.u, modes, 1, m, MODES, s)
    u1.fill_zero(s)
    nn.fourier_faces(u1, modes, 1, n, MODES, s)
    s.sync()
    t0 = time.perf_counter()
    net.forward()
    nn.trilinear(net.pred, u1, m, n, s)
    u2.copy_from(u1, s)
    rel_start = residual_norm(q, u1, s) / r0
    t_net = time.perf_counter() - t0
    sw, rel = jacobi_to_tolerance(u1, u2, n, r0, tol, check_every, max_sweeps, s)
    s.sync()
    out["learned"] = {"sweeps": sw, "seconds": time.perf_counter() - t0, "rel_residual": rel,
                      "start_rel_residual": rel_start, "network_seconds": t_net}
    # the network's input guess alone (transfinite blend of the faces)
    u1.fill_zero(s)
    nn.fourier_faces(u1, modes, 1, n, MODES, s)
    feat = DeviceArray((n ** 3, FEAT_C), np.uint16)
    base = DeviceArray((n, n, n), np.float32)
    nn.features(u1, feat, base, 1, n, FEAT_C, s)
    u1.copy_from(base, s)
    u2.copy_from(u1, s)
    rel_b = residual_norm(q, u1, s) / r0
    t0 = time.perf_counter()
    sw, rel = jacobi_to_tolerance(u1, u2, n, r0, tol, check_every, max_sweeps, s)
    s.sync()
    out["blend"] = {"sweeps": sw, "seconds": time.perf_counter() - t0, "rel_residual": rel,
                    "start_rel_residual": rel_b}
    out["time_ratio_learned_vs_plain"] = out["learned"]["seconds"] / out["plain"]["seconds"]
    if coarse_sweeps > 0:
        cm = DeviceArray((m, m, m), np.float32)
        cs = DeviceArray((m, m, m), np.float32)
        fm = DeviceArray((m ** 3, FEAT_C), np.uint16)
        for label in ("learned_coarse", "blend_coarse"):
            u1.fill_zero(s)
            nn.fourier_faces(u1, modes, 1, n, MODES, s)
            s.sync()
            t0 = time.perf_counter()
            if label == "learned_coarse":
                net.forward()
                cm.copy_from(net.pred, s)
            else:
                cm.fill_zero(s)
                nn.fourier_faces(cm, modes, 1, m, MODES, s)
                nn.features(cm, fm, cs, 1, m, FEAT_C, s)
                cm.copy_from(cs, s)
            coarse_refine(cm, cs, m, coarse_sweeps, s)
            nn.trilinear(cm, u1, m, n, s)
            u2.copy_from(u1, s)
            sw, rel = jacobi_to_tolerance(u1, u2, n, r0, tol, check_every, max_sweeps, s)
            s.sync()
            out[label] = {"sweeps": sw, "seconds": time.perf_counter() - t0, "rel_residual": rel,
                          "coarse_sweeps": coarse_sweeps}
        out["time_ratio_learned_coarse_vs_plain"] = out["learned_coarse"]["seconds"] / out["plain"]["seconds"]
        out["time_ratio_blend_coarse_vs_plain"] = out["blend_coarse"]["seconds"] / out["plain"]["seconds"]
        for t in (cm, cs, fm):
            t.free()
    for t in (modes, u1, u2, feat, base):
        t.free()
    return out


# ------------------------------------------------------------------ command line

def main(argv=None):
    """python -m thesis_amd.warmstart train|eval ...

    train: generate `--count` converged `--n`^3 solves with the batched
           Jacobi sweep, train the network, save the weights (.npz)
    eval:  compare plain zero-start JACOBI with the learned warm start on
           held-out 128^3 problems; prints one JSON line"""
    import argparse

    ap = argparse.ArgumentParser(prog="python -m thesis_amd.warmstart")
    ap.add_argument("mode", choices=["train", "eval"])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--sweeps", type=int, default=12000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0x5EED0005)
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      "tuned", "unet64.npz"))
    ap.add_argument("--out", default=None)
    ap.add_argument("--init", default=None, help="train: start from these weights (fine-tuning)")
    ap.add_argument("--fine", type=int, default=128)
    ap.add_argument("--problems", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=200)
    ap.add_argument("--coarse-sweeps", type=int, default=0,
                    help="eval: also time the prediction refined by this many (even) coarse sweeps")
    ap.add_argument("--smooth-sweeps", type=int, default=64)
    ap.add_argument("--smooth-weight", type=float, default=100.0)
    a = ap.parse_args(argv)
    s = Stream()
    if a.mode == "train":
        t0 = time.perf_counter()
        fields, _ = generate_solves(a.count, a.n, a.seed, a.sweeps, s)
        t_gen = time.perf_counter() - t0
        print(f"[warmstart] {a.count} solves of {a.n}^3 x {a.sweeps} sweeps: {t_gen:.1f} s", flush=True)
        net = UNet(a.batch, a.n, stream=s)
        if a.init:
            net.load(a.init)
        t0 = time.perf_counter()
        log = train(net, fields, a.count, a.steps, lr=a.lr, lr_final=a.lr / 10, log_every=100,
                    verbose=True, smooth_sweeps=a.smooth_sweeps, smooth_weight=a.smooth_weight)
        s.sync()
        t_train = time.perf_counter() - t0
        out = a.out or a.weights
        net.save(out)
        print(json.dumps({"generate_s": round(t_gen, 2), "train_s": round(t_train, 2),
                          "ms_per_step": round(1e3 * t_train / a.steps, 2), "loss": log,
                          "weights": out}), flush=True)
    else:
        with np.load(a.weights, allow_pickle=False) as z:
            state = {k: z[k] for k in z.files}
        results = []
        for i in range(a.problems):
            r = compare_warm_start(state, seed=a.seed + 1_000_000 + i, n=a.fine, m=a.n, tol=a.tol,
                                   check_every=a.check_every, stream=s, coarse_sweeps=a.coarse_sweeps)
            line = (f"[warmstart] problem {i}: plain {r['plain']['sweeps']} sweeps "
                    f"{r['plain']['seconds']:.3f} s, learned {r['learned']['sweeps']} sweeps "
                    f"{r['learned']['seconds']:.3f} s (ratio {r['time_ratio_learned_vs_plain']:.3f}), "
                    f"blend {r['blend']['sweeps']} sweeps")
            if a.coarse_sweeps > 0:
                line += (f"; +{a.coarse_sweeps} coarse sweeps: learned {r['learned_coarse']['sweeps']} "
                         f"sweeps (ratio {r['time_ratio_learned_coarse_vs_plain']:.3f}), blend "
                         f"{r['blend_coarse']['sweeps']} sweeps (ratio "
                         f"{r['time_ratio_blend_coarse_vs_plain']:.3f})")
            print(line, flush=True)
            results.append(r)
        ratios = [r["time_ratio_learned_vs_plain"] for r in results]
        doc = {"metric": "time-to-1e-6 residual, learned warm start vs plain zero start",
               "value": float(np.mean(ratios)), "unit": "ratio", "higher_is_better": False,
               "results": results}
        if a.coarse_sweeps > 0:
            doc["learned_coarse_mean_ratio"] = float(np.mean([r["time_ratio_learned_coarse_vs_plain"]
                                                              for r in results]))
            doc["blend_coarse_mean_ratio"] = float(np.mean([r["time_ratio_blend_coarse_vs_plain"]
                                                            for r in results]))
        print(json.dumps(doc), flush=True)


if __name__ == "__main__":
    main()
