"""ctypes bindings of the U-Net kernels (st_nn_* in thesis_amd/csrc/unet.hip).

Every function launches on the given stream and returns immediately; all
tensors are DeviceArrays (BF16 tensors are stored as uint16) or raw device
pointers.  Shapes are channels-last (N, D, H, W, C).
"""
from __future__ import annotations

import ctypes as C

from . import _check, _ptr, _sh, lib

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    P, I, LL, F, D = C.c_void_p, C.c_int, C.c_longlong, C.c_float, C.c_double
    sig = {
        "st_nn_kpad": ([I], I),
        "st_nn_conv3": ([P, P, P, P, P, I, I, I, I, I, I, I, I, P], I),
        "st_nn_conv3_wgrad": ([P, P, P, I, I, I, I, I, I, P], I),
        "st_nn_wt_transpose": ([P, P, I, I, P], I),
        "st_nn_f32_to_bf16": ([P, P, LL, P], I),
        "st_nn_bf16_to_f32": ([P, P, LL, P], I),
        "st_nn_relu_bwd": ([P, P, P, LL, P], I),
        "st_nn_bias_grad": ([P, P, LL, I, P], I),
        "st_nn_pool_fwd": ([P, P, I, I, I, I, I, P], I),
        "st_nn_pool_bwd": ([P, P, I, I, I, I, I, P], I),
        "st_nn_up_fwd": ([P, P, I, I, I, I, I, I, I, P], I),
        "st_nn_up_bwd": ([P, P, I, I, I, I, I, I, I, P], I),
        "st_nn_chan_copy": ([P, P, LL, I, I, I, I, I, I, I, P], I),
        "st_nn_head_fwd": ([P, P, P, P, LL, I, P], I),
        "st_nn_head_bwd": ([P, P, P, P, P, P, LL, I, P], I),
        "st_nn_mse": ([P, P, P, P, P, P, I, I, D, P], I),
        "st_nn_adam": ([P, P, P, P, P, LL, F, F, F, F, I, P], I),
        "st_nn_features": ([P, P, P, I, I, I, P], I),
        "st_nn_axpby": ([P, F, P, F, P, LL, P], I),
        "st_nn_sumsq": ([P, P, P, LL, D, P], I),
        "st_nn_trilinear": ([P, P, I, I, P], I),
        "st_nn_fourier_faces": ([P, P, I, I, I, P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _bound = True


def _p(x):
    if x is None:
        return None
    p = _ptr(x)
    return p


def kpad(ci: int) -> int:
    _bind()
    return lib().st_nn_kpad(ci)


def conv3(act, w, bias, out_bf, out_f32, shape, ci, co, relu=False, accumulate=False, stream=None):
    _bind()
    N, D, H, W = shape
    _check(lib().st_nn_conv3(_p(act), _p(w), _p(bias), _p(out_bf), _p(out_f32), N, D, H, W, ci, co,
                             int(relu), int(accumulate), _sh(stream)))


def conv3_wgrad(x, dy, dw, shape, ci, co, stream=None):
    _bind()
    N, D, H, W = shape
    _check(lib().st_nn_conv3_wgrad(_p(x), _p(dy), _p(dw), N, D, H, W, ci, co, _sh(stream)))


def wt_transpose(w, wt, ci, co, stream=None):
    _bind()
    _check(lib().st_nn_wt_transpose(_p(w), _p(wt), ci, co, _sh(stream)))


def f32_to_bf16(a, b, n, stream=None):
    _bind()
    _check(lib().st_nn_f32_to_bf16(_p(a), _p(b), n, _sh(stream)))


def bf16_to_f32(a, b, n, stream=None):
    _bind()
    _check(lib().st_nn_bf16_to_f32(_p(a), _p(b), n, _sh(stream)))


def relu_bwd(g, y, out, n, stream=None):
    _bind()
    _check(lib().st_nn_relu_bwd(_p(g), _p(y), _p(out), n, _sh(stream)))


def bias_grad(g, db, P, C_, stream=None):
    _bind()
    _check(lib().st_nn_bias_grad(_p(g), _p(db), P, C_, _sh(stream)))


def pool_fwd(x, y, shape_out, C_, stream=None):
    _bind()
    N, D, H, W = shape_out
    _check(lib().st_nn_pool_fwd(_p(x), _p(y), N, D, H, W, C_, _sh(stream)))


def pool_bwd(gout, gin, shape_out, C_, stream=None):
    _bind()
    N, D, H, W = shape_out
    _check(lib().st_nn_pool_bwd(_p(gout), _p(gin), N, D, H, W, C_, _sh(stream)))


def up_fwd(x, y, shape_in, C_, Ct, c0, stream=None):
    _bind()
    N, D, H, W = shape_in
    _check(lib().st_nn_up_fwd(_p(x), _p(y), N, D, H, W, C_, Ct, c0, _sh(stream)))


def up_bwd(gout, gin, shape_in, C_, Ct, c0, stream=None):
    _bind()
    N, D, H, W = shape_in
    _check(lib().st_nn_up_bwd(_p(gout), _p(gin), N, D, H, W, C_, Ct, c0, _sh(stream)))


def chan_copy(src, dst, P, C_, Cs, sc0, Cd, dc0, is_f32, accumulate=False, stream=None):
    _bind()
    _check(lib().st_nn_chan_copy(_p(src), _p(dst), P, C_, Cs, sc0, Cd, dc0, int(is_f32),
                                 int(accumulate), _sh(stream)))


def head_fwd(x, w, b, y, P, C_, stream=None):
    _bind()
    _check(lib().st_nn_head_fwd(_p(x), _p(w), _p(b), _p(y), P, C_, _sh(stream)))


def head_bwd(x, w, g, gx, dw, db, P, C_, stream=None):
    _bind()
    _check(lib().st_nn_head_bwd(_p(x), _p(w), _p(g), _p(gx), _p(dw), _p(db), P, C_, _sh(stream)))


def mse(base, y, target, g, loss, pred, N, n, scale, stream=None):
    _bind()
    _check(lib().st_nn_mse(_p(base), _p(y), _p(target), _p(g), _p(loss), _p(pred), N, n, scale,
                           _sh(stream)))


def adam(w, g, m, v, wbf, n, lr, b1, b2, eps, step, stream=None):
    _bind()
    _check(lib().st_nn_adam(_p(w), _p(g), _p(m), _p(v), _p(wbf), n, lr, b1, b2, eps, step, _sh(stream)))


def features(u, feat, base, N, n, C_, stream=None):
    _bind()
    _check(lib().st_nn_features(_p(u), _p(feat), _p(base), N, n, C_, _sh(stream)))


def trilinear(src, dst, m, M, stream=None):
    _bind()
    _check(lib().st_nn_trilinear(_p(src), _p(dst), m, M, _sh(stream)))


def fourier_faces(u, modes, N, n, M, stream=None):
    _bind()
    _check(lib().st_nn_fourier_faces(_p(u), _p(modes), N, n, M, _sh(stream)))


def axpby(out, a, x, b, y, n, stream=None):
    _bind()
    _check(lib().st_nn_axpby(_p(out), a, _p(x), b, _p(y), n, _sh(stream)))


def sumsq(x, g, loss, n, scale, stream=None):
    _bind()
    _check(lib().st_nn_sumsq(_p(x), _p(g), _p(loss), n, scale, _sh(stream)))
