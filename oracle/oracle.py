"""Python view of the CPU oracle (oracle/stencil_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by the cpu_baseline leg of ``bench.py``.  The product
path (``thesis_amd``) never imports it.

Every function mirrors a C entry point of ``libstencil_oracle.so``; arrays are
numpy, x fastest (``idx = i + j*nx + k*nx*ny``), shaped ``(nz, ny, nx)`` in 3D
and ``(ny, nx)`` in 2D.  Reference citations are in the C file's header.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

OP_THESIS = 0
OP_JACOBI = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libstencil_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle shared library (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        lng, i32, f32, f64, u64 = C.c_long, C.c_int, C.c_float, C.c_double, C.c_uint64
        L.or_uniform.argtypes = [u64, u64]
        L.or_uniform.restype = f32
        L.or_init_reference_f32.argtypes = [P, lng, lng, lng, i32]
        L.or_init_seeded_f32.argtypes = [P, lng, lng, lng, i32, i32, u64]
        L.or_init_seeded_f64.argtypes = [P, lng, lng, lng, i32, i32, u64]
        L.or_sweep_f32.argtypes = [P, P, lng, lng, lng, i32, i32, i32, f32, i32]
        L.or_sweep_f64.argtypes = [P, P, lng, lng, lng, i32, i32, i32, f64, i32]
        L.or_run_f32.argtypes = [P, P, lng, lng, lng, i32, i32, i32, f32, i32, i32]
        L.or_run_f32.restype = i32
        L.or_run_f64.argtypes = [P, P, lng, lng, lng, i32, i32, i32, f64, i32, i32]
        L.or_run_f64.restype = i32
        L.or_residual_norm_f32.argtypes = [P, lng, lng, lng, i32, i32, f64, i32]
        L.or_residual_norm_f32.restype = f64
        L.or_residual_norm_f64.argtypes = [P, lng, lng, lng, i32, i32, f64, i32]
        L.or_residual_norm_f64.restype = f64
        L.or_sse_f32.argtypes = [P, P, lng]
        L.or_sse_f32.restype = f64
        L.or_round16_array.argtypes = [P, P, lng, i32]
        L.or_widen16_array.argtypes = [P, P, lng, i32]
        L.or_mixed_run.argtypes = [P, P, P, P, P, lng, lng, lng, i32, i32, f32, i32, i32, i32, i32]
        L.or_mixed_run.restype = i32
        L.or_num_threads.restype = i32
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def shape_of(nx: int, ny: int, nz: int, dims: int):
    return (nz, ny, nx) if dims == 3 else (ny, nx)


def default_weight(dims: int, radius: int) -> float:
    """JACOBI neighbour weight 1/(2*D*R), rounded to FP32 as the GPU uses it."""
    return float(np.float32(1.0 / (2 * dims * radius)))


def uniform(seed: int, idx: int) -> float:
    return lib().or_uniform(seed, idx)


def init_reference(nx, ny, nz, dims, dtype=np.float32):
    """Reference initialisation (buffer 1): outer layer 1.0, rest 0."""
    u = np.empty(shape_of(nx, ny, nz, dims), np.float32)
    lib().or_init_reference_f32(_ptr(u), nx, ny, nz, dims)
    return u.astype(dtype) if dtype != np.float32 else u


def init_seeded(nx, ny, nz, dims, radius, seed, dtype=np.float32):
    """Seeded R-thick boundary shell from U(seed, idx); interior 0."""
    if dtype == np.float64:
        u = np.empty(shape_of(nx, ny, nz, dims), np.float64)
        lib().or_init_seeded_f64(_ptr(u), nx, ny, nz, dims, radius, seed)
    else:
        u = np.empty(shape_of(nx, ny, nz, dims), np.float32)
        lib().or_init_seeded_f32(_ptr(u), nx, ny, nz, dims, radius, seed)
    return u


def sweep(u_in: np.ndarray, u_out: np.ndarray, dims, radius, op, weight=None, nthreads=0):
    """One sweep of the interior from u_in into u_out (boundary of u_out untouched)."""
    nz = u_in.shape[0] if dims == 3 else 1
    ny, nx = u_in.shape[-2], u_in.shape[-1]
    w = default_weight(dims, radius) if weight is None else weight
    if u_in.dtype == np.float64:
        lib().or_sweep_f64(_ptr(u_in), _ptr(u_out), nx, ny, nz, dims, radius, op, w, nthreads)
    else:
        lib().or_sweep_f32(_ptr(u_in), _ptr(u_out), nx, ny, nz, dims, radius, op, w, nthreads)


def run(u1: np.ndarray, u2: np.ndarray, dims, radius, op, iterations, weight=None, nthreads=0):
    """`iterations` ping-pong sweeps in place; returns the array holding the result."""
    nz = u1.shape[0] if dims == 3 else 1
    ny, nx = u1.shape[-2], u1.shape[-1]
    if u1.dtype == np.float64:
        w = 1.0 / (2 * dims * radius) if weight is None else weight
        in_u2 = lib().or_run_f64(_ptr(u1), _ptr(u2), nx, ny, nz, dims, radius, op, w,
                                 iterations, nthreads)
    else:
        w = default_weight(dims, radius) if weight is None else weight
        in_u2 = lib().or_run_f32(_ptr(u1), _ptr(u2), nx, ny, nz, dims, radius, op, w,
                                 iterations, nthreads)
    return u2 if in_u2 else u1


def reference_run(nx, ny, nz, dims, radius, iterations, nthreads=0):
    """stencil_cpu() as the reference runs it: THESIS operator, ring init in
    buffer 1, buffer 2 zero-filled; returns the final field."""
    u1 = init_reference(nx, ny, nz, dims)
    u2 = np.zeros_like(u1)
    return run(u1, u2, dims, radius, OP_THESIS, iterations, nthreads=nthreads).copy()


def residual_norm(u: np.ndarray, dims, radius, weight=None, nthreads=0) -> float:
    """||w*S(u) - u||_2 over the interior with FP64 accumulation."""
    nz = u.shape[0] if dims == 3 else 1
    ny, nx = u.shape[-2], u.shape[-1]
    w = 1.0 / (2 * dims * radius) if weight is None else weight
    if u.dtype == np.float64:
        return lib().or_residual_norm_f64(_ptr(u), nx, ny, nz, dims, radius, w, nthreads)
    return lib().or_residual_norm_f32(_ptr(u), nx, ny, nz, dims, radius, w, nthreads)


def sse(a: np.ndarray, b: np.ndarray) -> float:
    """check_domain_errors: sum of squared FP32 differences, FP64 accumulation."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().or_sse_f32(_ptr(a), _ptr(b), a.size)


def round16(a: np.ndarray, is_f16: bool) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint16)
    lib().or_round16_array(_ptr(a), _ptr(out), a.size, int(is_f16))
    return out


def widen16(h: np.ndarray, is_f16: bool) -> np.ndarray:
    h = np.ascontiguousarray(h, np.uint16)
    out = np.empty(h.shape, np.float32)
    lib().or_widen16_array(_ptr(h), _ptr(out), h.size, int(is_f16))
    return out


def mixed_run(x0: np.ndarray, dims, radius, iterations, restart, is_f16, weight=None,
              nthreads=0):
    """16-bit-storage JACOBI with FP32 defect correction every `restart` sweeps.
    Returns (u_final_fp32, folds)."""
    x = np.ascontiguousarray(x0, np.float32).copy()
    nz = x.shape[0] if dims == 3 else 1
    ny, nx = x.shape[-2], x.shape[-1]
    e1 = np.empty(x.shape, np.uint16)
    e2 = np.empty(x.shape, np.uint16)
    d = np.empty(x.shape, np.uint16)
    u = np.empty(x.shape, np.float32)
    w = default_weight(dims, radius) if weight is None else weight
    folds = lib().or_mixed_run(_ptr(x), _ptr(e1), _ptr(e2), _ptr(d), _ptr(u), nx, ny, nz,
                               dims, radius, w, iterations, restart, int(is_f16), nthreads)
    return u, folds


def num_threads() -> int:
    return lib().or_num_threads()
