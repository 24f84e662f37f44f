"""stencil_dispatch_noalloc (the dispatch for callers that cannot allocate)
against stencil_dispatch, through the C ABI: bit for bit, on two calls in a
row, the second with a smaller grid, for the streaming form (AUTO at a grid
above the last-level cache, shell flags in scratch part 3) and the row form
(shell flag in part 2), with equal and with different shells."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(st):
    L = st.lib()
    f = L.stencil_dispatch_noalloc
    f.argtypes = [C.POINTER(st.StencilParams), C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    f.restype = C.c_int
    return f


def _run(st, p, init, fn):
    d1 = st.DeviceArray(p.shape, np.float32)
    d2 = st.DeviceArray(p.shape, np.float32)
    st.initialize_region(p, d1, d2, init, seed=0x5EED0042)
    if fn is None:
        in_u2, _ = st.dispatch_kernels(p, d1, d2)
    else:
        flag = C.c_int(-1)
        rc = fn(C.byref(p), d1.ptr, d2.ptr, None, C.byref(flag))
        if rc != 0:
            raise st.StencilError(rc, st.lib().stencil_last_error().decode())
        in_u2 = bool(flag.value)
    st.synchronize()
    out = (d2 if in_u2 else d1).numpy()
    d1.free()
    d2.free()
    return out, in_u2


@pytest.mark.parametrize("init", ["seeded", "reference"])
def test_noalloc_matches_dispatch(gpu, init):
    st = gpu
    fn = _noalloc(st)
    kind = st.INIT_SEEDED if init == "seeded" else st.INIT_REFERENCE
    op = st.OP_JACOBI if init == "seeded" else st.OP_THESIS
    calls = [
        # streaming form (fused passes; band fix-up with different shells)
        st.StencilParams.make(dim=336, radius=1, iterations=10, op=op, variant="auto", heuristic=1),
        st.StencilParams.make(nx=200, ny=90, nz=70, radius=1, iterations=9, op=op, variant="coop",
                              temporal_form=3, launch_per_pass=1, heuristic=1),
        # row form (whole-row stores with equal shells)
        st.StencilParams.make(nx=256, ny=120, nz=100, radius=2, iterations=5, op=op,
                              variant="smem_register", temporal_form=2, heuristic=1),
        st.StencilParams.make(nx=128, ny=40, nz=36, radius=4, iterations=3, op=op,
                              variant="smem_register", temporal_form=2, heuristic=1),
    ]
    assert st.launch_config(calls[0])["variant"] == st.V_COOP
    for p in calls:
        print("[case] noalloc", init, p.nx, p.ny, p.nz, p.radius, p.iterations, flush=True)
        want, in_a = _run(st, p, kind, None)
        got, in_b = _run(st, p, kind, fn)
        assert in_a == in_b == bool(p.iterations % 2)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (init, p.nx, p.ny, p.nz)
