"""MI355X (gfx950) iterative star-stencil sweep: Python host mirror of the C ABI.

The product is ``libthesis_stencil.so`` (HIP kernels + the C ABI declared in
``include/stencil_abi.h``).  This module binds it with ctypes and mirrors the
reference's host interface:

* :class:`StencilParams`  the knob surface (``-D DIM DIMENSIONS RADIUS
  ITERATIONS BLOCK_X/Y/Z COARSEN_X SMEM PADDED REGISTER COOP HEURISTIC
  SMEM_PAD``; mrfarstad/thesis Makefile:21-71, include/constants.h:1-78);
* :func:`dispatch_kernels`  ``dispatch_kernels(d_u1, d_u2)``
  (src/stencil_dispatch.cu:138-158);
* :func:`check_early_exit`  (include/stencil_utils.h:113-124), raising
  instead of exiting;
* :func:`initialize_region`  (include/stencil_initializer.h:6-52).

There is no CPU fallback: every entry point calls the HIP library and raises
:class:`StencilError` if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "StencilParams", "LaunchInfo", "StencilError", "DeviceArray", "Stream",
    "lib", "library_path", "dispatch_kernels", "sweep", "initialize_region",
    "residual_norm", "dispatch_mixed", "launch_config", "check_early_exit",
    "device_count", "device_name", "synchronize",
    "OP_THESIS", "OP_JACOBI", "V_BASE", "V_SMEM", "V_SMEM_PADDED", "V_SMEM_REGISTER",
    "V_COOP", "V_AUTO", "F32", "F64", "BF16", "F16", "INIT_REFERENCE", "INIT_SEEDED",
    "VARIANTS",
]

OP_THESIS, OP_JACOBI = 0, 1
V_BASE, V_SMEM, V_SMEM_PADDED, V_SMEM_REGISTER, V_COOP, V_AUTO = range(6)
F32, F64, BF16, F16 = range(4)
INIT_REFERENCE, INIT_SEEDED = 0, 1
VARIANTS = {"base": V_BASE, "smem": V_SMEM, "smem_padded": V_SMEM_PADDED,
            "smem_register": V_SMEM_REGISTER, "coop": V_COOP, "auto": V_AUTO}
STATUS = {0: "ok", 1: "invalid argument", 2: "LDS budget exceeded", 3: "launch failed",
          4: "out of memory", 5: "unsupported configuration", 6: "no usable device"}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libthesis_stencil.so"


def library_path() -> str:
    return os.path.join(_HERE, _LIB_NAME)


class StencilParams(C.Structure):
    """Mirror of ``st_params`` (include/stencil_abi.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dimensions", C.c_int), ("radius", C.c_int), ("iterations", C.c_int),
        ("block_x", C.c_int), ("block_y", C.c_int), ("block_z", C.c_int),
        ("coarsen_x", C.c_int), ("coarsen_y", C.c_int), ("smem_pad", C.c_int),
        ("variant", C.c_int), ("op", C.c_int), ("dtype", C.c_int),
        ("heuristic", C.c_int), ("temporal_steps", C.c_int), ("restart", C.c_int),
        ("launch_per_pass", C.c_int), ("temporal_form", C.c_int),
        ("weight", C.c_double),
    ]

    @classmethod
    def default(cls) -> "StencilParams":
        p = cls()
        lib().stencil_params_default(C.byref(p))
        return p

    @classmethod
    def make(cls, **kw) -> "StencilParams":
        """Defaults of include/constants.h overridden by keyword knobs.
        ``dim=N`` sets nx=ny=nz=N as the reference's DIM does."""
        p = cls.default()
        if "dim" in kw:
            d = kw.pop("dim")
            p.nx = p.ny = p.nz = d
        if isinstance(kw.get("variant"), str):
            kw["variant"] = VARIANTS[kw["variant"]]
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(f"unknown knob {k}")
            setattr(p, k, v)
        return p

    def copy(self, **kw) -> "StencilParams":
        q = StencilParams()
        C.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, VARIANTS[v] if k == "variant" and isinstance(v, str) else v)
        return q

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx) if self.dimensions == 3 else (self.ny, self.nx)

    @property
    def numpy_dtype(self):
        return {F32: np.float32, F64: np.float64, BF16: np.uint16, F16: np.uint16}[self.dtype]

    def __repr__(self):
        f = ", ".join(f"{n}={getattr(self, n)}" for n, _ in self._fields_)
        return f"StencilParams({f})"


class LaunchInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "variant", "block_x", "block_y", "block_z", "grid_x", "grid_y", "grid_z",
        "coarsen_x", "coarsen_y", "zchunk", "temporal_steps", "lds_bytes", "launches")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Problem(C.Structure):
    """Mirror of ``st_problem``."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("radius", C.c_int),
                ("iterations", C.c_int), ("op", C.c_int), ("weight", C.c_double),
                ("u1", C.c_void_p), ("u2", C.c_void_p)]


class TileRule(C.Structure):
    """Mirror of ``st_tile_rule``."""
    _fields_ = [("radius", C.c_int), ("max_cells", C.c_longlong), ("coarsen_y", C.c_int),
                ("zchunk", C.c_int)]


BATCH_GRAPH, BATCH_PERSISTENT, BATCH_SEQUENTIAL = 0, 1, 2


class StencilError(RuntimeError):
    def __init__(self, status: int, detail: str):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {detail}")


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise StencilError(6, f"{path} is missing; run __graft_entry__.build() first")
        L = C.CDLL(path)
        P, I, U64, D = C.c_void_p, C.c_int, C.c_uint64, C.c_double
        PP = C.POINTER(StencilParams)
        L.stencil_params_default.argtypes = [PP]
        L.stencil_params_default.restype = None
        L.stencil_validate.argtypes = [PP]
        L.stencil_launch_config.argtypes = [PP, C.POINTER(LaunchInfo)]
        L.stencil_sweep.argtypes = [PP, P, P, P]
        L.stencil_dispatch.argtypes = [PP, P, P, P, C.POINTER(I), C.POINTER(C.c_float)]
        L.stencil_init.argtypes = [PP, I, U64, P, P, P]
        L.stencil_residual_norm.argtypes = [PP, P, C.POINTER(D), P]
        L.stencil_dispatch_mixed.argtypes = [PP, P, P, P, P, P, P, C.POINTER(C.c_float)]
        L.stencil_status_str.argtypes = [I]
        L.stencil_status_str.restype = C.c_char_p
        L.stencil_last_error.argtypes = []
        L.stencil_last_error.restype = C.c_char_p
        L.st_device_count.argtypes = [C.POINTER(I)]
        L.st_device_name.argtypes = [I, C.c_char_p, I]
        L.st_malloc.argtypes = [C.POINTER(P), U64]
        L.st_free.argtypes = [P]
        L.st_memcpy_htod.argtypes = [P, P, U64, P]
        L.st_memcpy_dtoh.argtypes = [P, P, U64, P]
        L.st_memcpy_dtod.argtypes = [P, P, U64, P]
        L.st_memset.argtypes = [P, I, U64, P]
        L.st_stream_create.argtypes = [C.POINTER(P)]
        L.st_stream_destroy.argtypes = [P]
        L.st_stream_sync.argtypes = [P]
        L.st_device_sync.argtypes = []
        L.st_copy_f32.argtypes = [P, P, U64, P]
        L.stencil_round16.argtypes = [P, P, U64, I, P]
        L.stencil_widen16.argtypes = [P, P, U64, I, P]
        L.stencil_dispatch_batched.argtypes = [C.POINTER(Problem), I, C.POINTER(TileRule), I, I, P,
                                               C.POINTER(C.c_float)]
        for name in ("stencil_validate", "stencil_launch_config", "stencil_sweep",
                     "stencil_dispatch", "stencil_init", "stencil_residual_norm",
                     "stencil_dispatch_mixed", "st_device_count", "st_device_name", "st_malloc",
                     "st_free", "st_memcpy_htod", "st_memcpy_dtoh", "st_memcpy_dtod",
                     "st_memset", "st_stream_create", "st_stream_destroy", "st_stream_sync",
                     "st_device_sync", "st_copy_f32", "stencil_round16", "stencil_widen16",
                     "stencil_dispatch_batched"):
            getattr(L, name).restype = I
        _lib = L
    return _lib


def _check(status: int):
    if status != 0:
        raise StencilError(status, lib().stencil_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    lib().st_device_count(C.byref(n))
    return n.value


def device_name(device: int = 0) -> str:
    buf = C.create_string_buffer(256)
    _check(lib().st_device_name(device, buf, 256))
    return buf.value.decode()


def synchronize():
    _check(lib().st_device_sync())


def copy_f32(dst, src, n: int, stream=None):
    """Streaming-copy probe kernel (n floats)."""
    _check(lib().st_copy_f32(_ptr(dst), _ptr(src), n, _sh(stream)))


def round16(d_in, d_out, n: int, dtype: int, stream=None):
    """d_out (uint16) <- round-to-nearest-even of n floats to BF16 or F16."""
    _check(lib().stencil_round16(_ptr(d_in), _ptr(d_out), n, dtype, _sh(stream)))


def widen16(d_in, d_out, n: int, dtype: int, stream=None):
    """d_out (float32) <- exact widening of n BF16 or F16 values."""
    _check(lib().stencil_widen16(_ptr(d_in), _ptr(d_out), n, dtype, _sh(stream)))


class Stream:
    """An owned hipStream_t."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().st_stream_create(C.byref(h)))
        self.handle = h

    def sync(self):
        _check(lib().st_stream_sync(self.handle))

    def close(self):
        if self.handle:
            lib().st_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sh(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class DeviceArray:
    """A device allocation with a numpy shape/dtype (x fastest)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().st_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().st_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, _sh(stream)))

    def numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().st_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, _sh(stream)))
        return out

    def fill_zero(self, stream=None):
        _check(lib().st_memset(self.ptr, 0, self.nbytes, _sh(stream)))

    def copy_from(self, other: "DeviceArray", stream=None):
        assert other.nbytes == self.nbytes
        _check(lib().st_memcpy_dtod(self.ptr, other.ptr, self.nbytes, _sh(stream)))

    def free(self):
        if self.ptr:
            lib().st_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(params: StencilParams):
    return C.byref(params)


def _ptr(x):
    return x.ptr if isinstance(x, DeviceArray) else x


def check_early_exit(params: StencilParams):
    """Validate knobs; raises StencilError where the reference exits 1."""
    _check(lib().stencil_validate(_p(params)))


def launch_config(params: StencilParams) -> dict:
    info = LaunchInfo()
    _check(lib().stencil_launch_config(_p(params), C.byref(info)))
    return info.as_dict()


def sweep(params: StencilParams, d_in, d_out, stream=None):
    """One sweep d_out <- S(d_in) (one get_kernel() launch)."""
    _check(lib().stencil_sweep(_p(params), _ptr(d_in), _ptr(d_out), _sh(stream)))


def dispatch_kernels(params: StencilParams, d_u1, d_u2, stream=None, timed: bool = False):
    """``params.iterations`` sweeps ping-ponging between d_u1 and d_u2.
    Returns (result_in_u2, elapsed_ms or None)."""
    in_u2 = C.c_int(0)
    ms = C.c_float(0.0)
    _check(lib().stencil_dispatch(_p(params), _ptr(d_u1), _ptr(d_u2), _sh(stream),
                                  C.byref(in_u2), C.byref(ms) if timed else None))
    return bool(in_u2.value), (ms.value if timed else None)


def initialize_region(params: StencilParams, d_u1, d_u2, kind: int = INIT_REFERENCE,
                      seed: int = 0, stream=None):
    _check(lib().stencil_init(_p(params), kind, seed, _ptr(d_u1), _ptr(d_u2), _sh(stream)))


def residual_norm(params: StencilParams, d_u, stream=None) -> float:
    out = C.c_double(0.0)
    _check(lib().stencil_residual_norm(_p(params), _ptr(d_u), C.byref(out), _sh(stream)))
    return out.value


def dispatch_mixed(params: StencilParams, d_x, d_e1, d_e2, d_d, d_u, stream=None,
                   timed: bool = False):
    ms = C.c_float(0.0)
    _check(lib().stencil_dispatch_mixed(_p(params), _ptr(d_x), _ptr(d_e1), _ptr(d_e2), _ptr(d_d),
                                        _ptr(d_u), _sh(stream), C.byref(ms) if timed else None))
    return ms.value if timed else None


def dispatch_batched(problems, rules=(), mode: int = BATCH_GRAPH, stream=None, timed: bool = False):
    """Advance a batch of independent grids.  `problems` is a sequence of
    Problem (or dicts with its fields, buffers as DeviceArray or pointers);
    `rules` a sequence of TileRule (or dicts)."""
    n = len(problems)
    parr = (Problem * max(n, 1))()
    for i, q in enumerate(problems):
        if isinstance(q, Problem):
            parr[i] = q
        else:
            parr[i] = Problem(q["nx"], q["ny"], q["nz"], q.get("radius", 1), q.get("iterations", 1),
                              q.get("op", OP_JACOBI), q.get("weight", 0.0), _vp(q["u1"]), _vp(q["u2"]))
    rl = list(rules)
    rarr = (TileRule * max(len(rl), 1))()
    for i, r in enumerate(rl):
        rarr[i] = r if isinstance(r, TileRule) else TileRule(r["radius"], r["max_cells"],
                                                             r["coarsen_y"], r["zchunk"])
    ms = C.c_float(0.0)
    _check(lib().stencil_dispatch_batched(parr, n, rarr, len(rl), mode, _sh(stream),
                                          C.byref(ms) if timed else None))
    return ms.value if timed else None


def _vp(x):
    p = _ptr(x)
    return p.value if isinstance(p, C.c_void_p) else p
