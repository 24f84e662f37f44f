"""Batched sweeps of uneven grids: every mode against the CPU oracle per
problem (bit-exact), and the three modes against each other on the full
512-problem batch."""
import hashlib

import numpy as np
import pytest

from thesis_amd import workloads

pytestmark = pytest.mark.gpu


def _alloc(st, oracle, shapes, its, ops, seed, stream):
    probs, host, bufs = [], [], []
    for i, (nx, ny, nz, r) in enumerate(shapes):
        u1 = oracle.init_seeded(nx, ny, nz, 3, r, seed + i)
        if ops[i] == 0:  # THESIS: random interior, buffer 2 shell zero
            rng = np.random.default_rng(i)
            u1 = np.where(u1 == 0, rng.random(u1.shape), u1).astype(np.float32)
            u2 = np.zeros_like(u1)
        else:
            u2 = u1.copy()
        d1 = st.DeviceArray.from_numpy(u1, stream)
        d2 = st.DeviceArray.from_numpy(u2, stream)
        probs.append(dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=its[i], op=ops[i], u1=d1, u2=d2))
        host.append((u1, u2))
        bufs += [d1, d2]
    return probs, host, bufs


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batched_small_vs_oracle(gpu, oracle, mode):
    st = gpu
    stream = st.Stream()
    shapes = workloads.batched_shapes(24, seed=77, lo=16, hi=48)
    its = [4 + (i % 5) for i in range(24)]
    ops = [i % 2 for i in range(24)]
    print(f"[case] batched mode {mode}", flush=True)
    probs, host, bufs = _alloc(st, oracle, shapes, its, ops, 1000, stream)
    rules = [dict(radius=1, max_cells=20000, coarsen_y=4, zchunk=8),
             dict(radius=2, max_cells=10 ** 9, coarsen_y=2, zchunk=0)]
    st.dispatch_batched(probs, rules, mode, stream=stream)
    stream.sync()
    for i, (q, (u1, u2)) in enumerate(zip(probs, host)):
        want = oracle.run(u1.copy(), u2.copy(), 3, q["radius"], q["op"], q["iterations"])
        got = (q["u2"] if q["iterations"] % 2 else q["u1"]).numpy(stream)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (i, shapes[i], q["op"])


def test_batched_full_modes_agree(gpu, oracle):
    st = gpu
    stream = st.Stream()
    shapes = workloads.batched_shapes(512)
    n = len(shapes)
    assert 5e7 < workloads.batched_cells(shapes) < 2e8
    its = [50] * n
    ops = [1] * n
    digests = {}
    for mode in (0, 1, 2):
        print(f"[case] full batch mode {mode}", flush=True)
        probs = []
        bufs = []
        for i, (nx, ny, nz, r) in enumerate(shapes):
            p = st.StencilParams.make(nx=nx, ny=ny, nz=nz, dimensions=3, radius=r, op=1)
            d1 = st.DeviceArray(p.shape, np.float32)
            d2 = st.DeviceArray(p.shape, np.float32)
            st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=0x5EED0004 + i, stream=stream)
            probs.append(dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=50, op=1, u1=d1, u2=d2))
            bufs += [d1, d2]
        ms = st.dispatch_batched(probs, [], mode, stream=stream, timed=True)
        stream.sync()
        h = hashlib.sha256()
        for q in probs:
            h.update(q["u1"].numpy(stream).tobytes())
        digests[mode] = h.hexdigest()
        print(f"[case] mode {mode}: {ms:.2f} ms", flush=True)
        for b in bufs:
            b.free()
    assert len(set(digests.values())) == 1, digests
    # spot-check one problem of the batch against the oracle
    nx, ny, nz, r = shapes[3]
    u = oracle.init_seeded(nx, ny, nz, 3, r, 0x5EED0004 + 3)
    want = oracle.run(u.copy(), u.copy(), 3, r, 1, 50)
    p = st.StencilParams.make(nx=nx, ny=ny, nz=nz, dimensions=3, radius=r, op=1)
    d1 = st.DeviceArray(p.shape, np.float32)
    d2 = st.DeviceArray(p.shape, np.float32)
    st.initialize_region(p, d1, d2, st.INIT_SEEDED, seed=0x5EED0004 + 3, stream=stream)
    st.dispatch_batched([dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=50, op=1, u1=d1, u2=d2)],
                        [], 0, stream=stream)
    assert np.array_equal(d1.numpy(stream).view(np.uint32), want.view(np.uint32))


_PLANE_SHAPES = [(3, 3, 3, 1), (5, 40, 7, 2), (130, 3, 9, 1), (300, 50, 6, 1), (700, 9, 5, 2), (64, 64, 64, 2),
                 (17, 200, 11, 1)]


@pytest.mark.parametrize("shapes,op", [
    # plane-chunk form: minimal grids, one-row frames (nx = 700, R = 2), many
    # frames, frames not dividing ny, chunks not dividing nz.  THESIS problems
    # start with buffer 2 zero (different shells: interior-only stores),
    # JACOBI ones with equal shells (whole-row stores)
    (_PLANE_SHAPES, 0),
    (_PLANE_SHAPES, 1),
    # mixed operators: the whole batch takes the march form
    (_PLANE_SHAPES, None),
    # one problem too wide for the plane form: the whole batch takes the march form
    ([(1400, 5, 5, 1), (20, 18, 16, 2), (33, 31, 29, 1)], None),
], ids=["plane-thesis", "plane-jacobi", "mixed-ops", "march"])
def test_batched_forms_edge_shapes(gpu, oracle, shapes, op):
    st = gpu
    stream = st.Stream()
    n = len(shapes)
    its = [3 + (i % 4) for i in range(n)]
    ops = [i % 2 if op is None else op for i in range(n)]
    for mode in (0, 1):
        print(f"[case] batched edge shapes mode {mode} op {op} {shapes}", flush=True)
        probs, host, bufs = _alloc(st, oracle, shapes, its, ops, 2000, stream)
        rules = [dict(radius=1, max_cells=10 ** 9, coarsen_y=0, zchunk=4),
                 dict(radius=2, max_cells=10 ** 9, coarsen_y=0, zchunk=0)]
        st.dispatch_batched(probs, rules, mode, stream=stream)
        stream.sync()
        for i, (q, (u1, u2)) in enumerate(zip(probs, host)):
            a, b = u1.copy(), u2.copy()
            oracle.run(a, b, 3, q["radius"], q["op"], q["iterations"])
            res = 1 if q["iterations"] % 2 else 0
            bufs_got = (q["u1"].numpy(stream), q["u2"].numpy(stream))
            bufs_want = (a, b)
            # the result buffer bit for bit, shells included
            assert np.array_equal(bufs_got[res].view(np.uint32), bufs_want[res].view(np.uint32)), \
                (mode, i, shapes[i], q["op"])
            # the other buffer: the previous sweep, except where the graph
            # or persistent mode fused two sweeps per pass (plane form, equal
            # shells, 4 sweeps or more), which leaves an intermediate state
            # in its interior; no shell cell may change either way
            plane = all((1 + 2 * r_) * nx_ <= 4096 and nx_ <= 4 * 512 // r_ for nx_, _, _, r_ in shapes) \
                and len(set(ops)) == 1
            fused = plane and np.array_equal(u1, u2) and q["iterations"] >= 4
            other_got, other_want = bufs_got[1 - res], bufs_want[1 - res]
            if fused:
                shell = np.ones(other_got.shape, bool)
                r = q["radius"]
                shell[r:-r, r:-r, r:-r] = False
                other_got, other_want = other_got[shell], other_want[shell]
            assert np.array_equal(other_got.view(np.uint32), other_want.view(np.uint32)), \
                (mode, i, shapes[i], q["op"])
        for b in bufs:
            b.free()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("r", [1, 2])
def test_batched_fused_passes(gpu, oracle, r, mode):
    """Graph mode (one launch per pass) and persistent mode (one cooperative
    launch, captured in a graph) with two sweeps fused per pass: problems with equal shells
    (every sweep count 1..9, so 0, 2 or 4 fused passes then 1-3 single
    sweeps) beside problems whose buffer 2 is zero (different shells, single
    sweeps only) in one plane-form batch; every result buffer bit-exact with
    the oracle, frames of one or several rows, chunks not dividing nz."""
    st = gpu
    stream = st.Stream()
    shapes = [(128, 40, 30), (64, 70, 23), (33, 17, 19), (17, 130, 9), (96, 9, 70), (200, 11, 8),
              (40, 40, 40), (128, 128, 12), (20, 20, 20)]
    probs, host, bufs = [], [], []
    for i, (nx, ny, nz) in enumerate(shapes):
        for eq in (True, False):
            u1 = oracle.init_seeded(nx, ny, nz, 3, r, 3000 + 2 * i + eq)
            rng = np.random.default_rng(i)
            u1 = np.where(u1 == 0, rng.random(u1.shape), u1).astype(np.float32)
            u2 = u1.copy() if eq else np.zeros_like(u1)
            its = 1 + (i + (0 if eq else 4)) % 9
            d1 = st.DeviceArray.from_numpy(u1, stream)
            d2 = st.DeviceArray.from_numpy(u2, stream)
            probs.append(dict(nx=nx, ny=ny, nz=nz, radius=r, iterations=its, op=1, u1=d1, u2=d2))
            host.append((u1, u2))
            bufs += [d1, d2]
    print(f"[case] batched fused passes r={r} mode={mode}", flush=True)
    rules = [dict(radius=r, max_cells=10 ** 9, coarsen_y=0, zchunk=13)]

    def check(rounds, only_even):
        for i, (q, (u1, u2)) in enumerate(zip(probs, host)):
            its = rounds * q["iterations"]
            if only_even and q["iterations"] % 2:
                continue
            want = oracle.run(u1.copy(), u2.copy(), 3, r, 1, its)
            got = (q["u2"] if its % 2 else q["u1"]).numpy(stream)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (rounds, i, q["nx"], q["ny"], q["nz"],
                                                                                q["iterations"])

    st.dispatch_batched(probs, rules, mode, stream=stream)
    stream.sync()
    check(1, False)
    # a second dispatch replays the cached graph; it continues from the
    # results where they are in u1 (an even sweep count)
    st.dispatch_batched(probs, rules, mode, stream=stream)
    stream.sync()
    check(2, True)
    for b in bufs:
        b.free()


def test_batched_persistent_null_stream(gpu, oracle):
    """The persistent mode on the null stream (no capture there): the
    cooperative launch runs directly, fused passes included."""
    st = gpu
    shapes = [(40, 30, 20, 1), (33, 17, 19, 2), (64, 70, 23, 1)]
    its = [6, 5, 8]
    probs, host, bufs = _alloc(st, oracle, shapes, its, [1, 1, 1], 4000, None)
    print("[case] batched persistent on the null stream", flush=True)
    st.dispatch_batched(probs, [], 1)
    st.synchronize()
    for i, (q, (u1, u2)) in enumerate(zip(probs, host)):
        want = oracle.run(u1.copy(), u2.copy(), 3, q["radius"], 1, q["iterations"])
        got = (q["u2"] if q["iterations"] % 2 else q["u1"]).numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (i, shapes[i])
    for b in bufs:
        b.free()
