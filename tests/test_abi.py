"""Host-side checks of the C ABI that need no GPU: the library loads, it
exports every function include/stencil_abi.h declares, and knob validation
(the check_early_exit analogue) accepts and rejects what it should."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.append(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("stencil_dispatch", "stencil_sweep", "stencil_validate", "stencil_launch_config",
                 "stencil_init", "stencil_residual_norm", "stencil_dispatch_mixed",
                 "stencil_status_str", "stencil_params_default"):
        assert must in names


def test_library_exports_every_declared_symbol(st):
    L = C.CDLL(st.library_path())
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_params_layout_matches_header(st):
    # st_params: 20 ints then a double -> 88 bytes
    assert C.sizeof(st.StencilParams) == 88
    p = st.StencilParams.default()
    assert (p.nx, p.ny, p.nz, p.dimensions, p.radius, p.iterations) == (256, 256, 256, 3, 1, 64)
    assert (p.block_x, p.block_y, p.block_z, p.coarsen_x) == (32, 2, 2, 4)
    assert p.variant == st.V_BASE and p.op == st.OP_THESIS and p.dtype == st.F32


def test_status_strings(st):
    L = st.lib()
    assert L.stencil_status_str(0) == b"ok"
    assert L.stencil_status_str(2) == b"LDS budget exceeded"


def _raises(st, status, **kw):
    with pytest.raises(st.StencilError) as e:
        st.check_early_exit(st.StencilParams.make(**kw))
    assert e.value.status == status, str(e.value)


def test_check_early_exit_rules(st):
    # include/stencil_utils.h:113-124: padded needs bx, by (bz in 3D) >= R;
    # register needs bx (by in 3D) >= R.
    ok = dict(dim=64, radius=4, iterations=2)
    _raises(st, 1, variant="smem_padded", block_x=2, block_y=8, block_z=8, **ok)
    _raises(st, 1, variant="smem_padded", block_x=8, block_y=2, block_z=8, **ok)
    _raises(st, 1, variant="smem_padded", block_x=8, block_y=8, block_z=2, **ok)
    st.check_early_exit(st.StencilParams.make(variant="smem_padded", block_x=8, block_y=8,
                                              block_z=2, dimensions=2, **ok))
    _raises(st, 1, variant="smem_register", block_x=2, block_y=8, block_z=8, **ok)
    _raises(st, 1, variant="smem_register", block_x=64, block_y=2, block_z=8, **ok)
    st.check_early_exit(st.StencilParams.make(variant="smem_register", block_x=64, block_y=4,
                                              block_z=16, **ok))
    # plain smem and base have no radius rule
    st.check_early_exit(st.StencilParams.make(variant="smem", block_x=2, block_y=2, block_z=2, **ok))


def test_invalid_knobs(st):
    _raises(st, 1, dim=2, radius=1)               # extent must exceed 2R
    _raises(st, 1, dimensions=4)
    _raises(st, 1, radius=0)
    _raises(st, 1, iterations=-1)
    _raises(st, 1, block_x=64, block_y=32, block_z=1, variant="base")  # 2048 threads
    _raises(st, 1, dtype=st.BF16)                # 16-bit storage only through the mixed path
    _raises(st, 2, variant="smem_padded", block_x=16, block_y=8, block_z=8, radius=8, dim=64,
            coarsen_x=4)                         # LDS over 160 KiB
    _raises(st, 5, variant="smem_register", radius=17, dim=64, heuristic=1)
    _raises(st, 5, variant="coop", radius=5, dim=64, heuristic=1, temporal_form=1)  # LDS rings: radius 1..4
    _raises(st, 5, variant="coop", dimensions=2, dim=64, radius=2, heuristic=1,
            temporal_form=3)                                                   # 2D streaming: radius 1
    _raises(st, 5, variant="coop", dimensions=2, dim=60, heuristic=1,
            temporal_form=3)                                                   # nx not a multiple of 8


def test_launch_config_shapes(st):
    base = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100, variant="base",
                                                  block_x=32, block_y=8, block_z=4, coarsen_x=1))
    assert (base["grid_x"], base["grid_y"], base["grid_z"]) == (16, 64, 128)  # 512/32, 512/8, 512/4
    assert base["launches"] == 100 and base["lds_bytes"] == 0
    pad = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100, variant="smem_padded",
                                                 block_x=32, block_y=8, block_z=4, coarsen_x=1))
    assert pad["lds_bytes"] == (32 + 2) * (8 + 2) * (4 + 2) * 4  # src/stencil_dispatch.cu:57
    sm = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100, variant="smem",
                                                block_x=32, block_y=8, block_z=4, coarsen_x=4))
    assert sm["lds_bytes"] == 32 * 4 * 8 * 4 * 4 and sm["grid_x"] == 4
    reg = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100,
                                                 variant="smem_register", heuristic=1, temporal_form=1))
    assert reg["block_x"] % 64 == 0 and reg["lds_bytes"] == 0 and reg["zchunk"] >= 16
    # the library's SMEM_REGISTER choice: row streaming, whole rows, two LDS
    # planes of (rows + 2R) rows of 8 cells per lane
    rows = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100,
                                                  variant="smem_register", heuristic=1))
    assert rows["block_x"] == 512 and rows["coarsen_x"] == 8
    assert rows["lds_bytes"] == 2 * (rows["block_y"] + 2) * 512 * 4
    coop = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=100, variant="coop",
                                                  heuristic=1, op=st.OP_JACOBI))
    assert coop["temporal_steps"] >= 1 and coop["lds_bytes"] <= 160 * 1024
    assert coop["launches"] == 1


def test_heuristic_keeps_coarsen_x(st):
    """HEURISTIC picks the block shape; COARSEN_X still selects the unrolled
    kernels (src/stencil_dispatch.cu:14-42, :148): the x grid shrinks by it."""
    for variant in ("base", "smem", "smem_padded"):
        for cx in (1, 2, 4, 8):
            cfg = st.launch_config(st.StencilParams.make(dim=512, radius=1, iterations=4, variant=variant,
                                                         heuristic=1, coarsen_x=cx))
            assert cfg["coarsen_x"] == cx, (variant, cx)
            assert cfg["grid_x"] == -(-512 // (cfg["block_x"] * cx)), (variant, cx)
    # the padded tile halves its coarsening until it fits in LDS
    big = st.launch_config(st.StencilParams.make(dim=256, radius=8, iterations=4, variant="smem_padded",
                                                 heuristic=1, coarsen_x=8))
    assert big["lds_bytes"] <= 160 * 1024 and big["coarsen_x"] < 8
    # the library's choice for small grids runs one cell per thread
    auto = st.launch_config(st.StencilParams.make(dim=64, radius=1, iterations=4, variant="auto"))
    assert st.VARIANTS["base"] == auto["variant"] and auto["coarsen_x"] == 1


def test_coop_row_form_config(st):
    """COOP at radius > 1 runs the row kernel with a grid barrier after each
    sweep: one launch for every sweep, or one per sweep with
    launch_per_pass."""
    for r in (2, 3, 4):
        cfg = st.launch_config(st.StencilParams.make(dim=384, radius=r, iterations=10, variant="coop",
                                                     heuristic=1))
        assert cfg["launches"] == 1 and cfg["temporal_steps"] == 1 and cfg["block_x"] == 384
        per = st.launch_config(st.StencilParams.make(dim=384, radius=r, iterations=10, variant="coop",
                                                     heuristic=1, launch_per_pass=1))
        assert per["launches"] == 10
    f4 = st.launch_config(st.StencilParams.make(dim=128, radius=1, iterations=7, variant="coop",
                                                temporal_form=4, heuristic=1))
    assert f4["launches"] == 1 and f4["temporal_steps"] == 1
    _raises(st, 5, variant="coop", temporal_form=4, nx=62, ny=40, nz=40, heuristic=1)
    _raises(st, 1, variant="coop", temporal_form=6, dim=64, heuristic=1)


def test_streaming_form_plans(st):
    """The streaming COOP form's frame and chunk plans (host logic, no GPU):
    the 512^3 benchmark's 8-wave frames of 16 rows (12 stored) in 103-plane
    chunks; over 512 columns two waves of 8 cells per row and 12-row frames
    (8 stored), the chunk count searched when the LDS caps the chunk; one
    wave of 16 cells per row on request; LDS within 160 KiB less the
    kernel's own 256 bytes."""
    def plan(n, **kw):
        p = st.StencilParams.make(dim=n, radius=1, iterations=100, op=st.OP_JACOBI, **kw)
        return st.launch_config(p)
    c = plan(512, variant="auto", heuristic=1)
    assert c["variant"] == st.VARIANTS["coop"]
    assert (c["coarsen_x"], c["coarsen_y"], c["block_y"], c["zchunk"]) == (8, 2, 12, 103), c
    for n, chunk in ((1024, 256), (896, 224), (768, 154), (1000, 250)):
        c = plan(n, variant="auto", heuristic=1)
        assert c["variant"] == st.VARIANTS["coop"], (n, c)
        assert (c["coarsen_x"], c["coarsen_y"], c["block_y"], c["zchunk"]) == (8, 2, 8, chunk), (n, c)
        assert c["lds_bytes"] <= 160 * 1024 - 256, (n, c)
    one = plan(1024, variant="coop", temporal_form=3, heuristic=0, block_x=1024, block_y=8, block_z=64,
               coarsen_y=1)
    assert (one["coarsen_x"], one["coarsen_y"], one["block_y"]) == (16, 1, 4), one
    # a chunk whose LDS slice does not fit is refused up front (ST_ELDS)
    with pytest.raises(st.StencilError) as e:
        plan(1024, variant="coop", temporal_form=3, heuristic=0, block_x=1024, block_y=6, block_z=400)
    assert e.value.status == 2, str(e.value)


def test_coop_grid_stride_plans(st):
    """COOP at radius 5 and up (3D) and on 2D grids the streaming form does
    not take runs the grid-stride form (the reference's coop kernel): one
    sweep per grid barrier, one cooperative launch or one launch per sweep;
    the LDS and register forms refuse radius 5 and up (host logic, no GPU)."""
    def cfg(**kw):
        kw.setdefault("variant", "coop")
        kw.setdefault("heuristic", 1)
        return st.launch_config(st.StencilParams.make(**kw))
    for r in (5, 8, 16, 32, 128):
        c = cfg(dim=300, radius=r, iterations=3)
        assert (c["variant"], c["temporal_steps"], c["launches"], c["lds_bytes"]) == \
            (st.VARIANTS["coop"], 1, 1, 0), (r, c)
    assert cfg(dim=40, radius=8, iterations=3, launch_per_pass=1)["launches"] == 3
    assert cfg(dim=40, radius=2, iterations=4, temporal_form=5)["temporal_steps"] == 1
    # 2D: nx not a multiple of 8, FP64 or radius 2 leave the streaming form
    for kw in (dict(nx=131, ny=77), dict(nx=128, ny=77, dtype=st.F64), dict(nx=128, ny=77, radius=2)):
        kw.setdefault("radius", 1)
        c = cfg(nz=1, dimensions=2, iterations=10, **kw)
        assert (c["temporal_steps"], c["launches"]) == (1, 1), (kw, c)
    assert cfg(nx=128, ny=77, nz=1, dimensions=2, radius=1, iterations=10)["temporal_steps"] > 1
    for form in (1, 2, 3, 4):
        with pytest.raises(st.StencilError) as e:
            cfg(dim=40, radius=8, iterations=3, temporal_form=form)
        assert "grid-stride" in str(e.value) or e.value.status != 0, str(e.value)
    with pytest.raises(st.StencilError):
        cfg(dim=40, radius=1, iterations=3, temporal_form=6)


def test_stream2d_schedules(st):
    """The 2D streaming form's pass schedule (host logic, no GPU): passes of
    six and four fused sweeps, an even number of them (so the state ends
    where single sweeps leave it), as few as possible and then as deep as
    possible; under 8 sweeps AUTO keeps a single-sweep kernel."""
    def cfg(its, **kw):
        kw.setdefault("variant", "auto")
        p = st.StencilParams.make(nx=32768, ny=32768, nz=1, dimensions=2, radius=1, iterations=its,
                                  op=st.OP_JACOBI, heuristic=1, **kw)
        return st.launch_config(p)
    # 8 = 4 + 4, 12 = 6 + 6, 64 = 8 x 6 + 4 x 4, 24 = 4 x 6
    for its, launches in ((8, 2), (12, 2), (64, 12), (24, 4)):
        c = cfg(its)
        assert c["variant"] == st.VARIANTS["coop"] and c["launches"] == launches, (its, c)
        assert (c["block_x"], c["coarsen_x"]) == (512, 8), c
    assert cfg(7)["variant"] != st.VARIANTS["coop"]
    # passes of four only on request: 64 = 16 x 4
    assert cfg(64, variant="coop", temporal_steps=4)["launches"] == 16
