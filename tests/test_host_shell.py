"""The 3-D shell of the tiled path (every point outside the tiles' box: faces, edges, boundary
rows), evaluated strip by strip with the device kernel's per-strip code compiled for the CPU
(tests/hostcheck: dop_strip.hpp with scalar loads), must equal the generic per-point
evaluator bit for bit at every shell point, and the strips must cover exactly the shell."""
import ctypes as C
import os

import numpy as np
import pytest

import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import _bcs_of, make_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC_PATH = os.environ.get("DOP_HOSTCHECK_LIBRARY") or os.path.join(ROOT, "tests", "hostcheck", "_build",
                                                                "libdop_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC_PATH):
        pytest.skip("tests/hostcheck not built")
    L = C.CDLL(HC_PATH)
    L.hc_apply.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p, C.c_void_p, C.c_int]
    L.hc_shell.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p, C.c_void_p, C.c_int]
    L.hc_shell.restype = C.c_int64
    L.hc_shell_kinds.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p]
    L.hc_last_error.restype = C.c_char_p
    return L


def kinds(hc, d):
    pts = np.zeros(4, dtype=np.int64)
    assert hc.hc_shell_kinds(C.byref(d), pts.ctypes.data) > 0
    return pts


STORAGE = {"f64": (_abi.F64, np.float64, np.uint64), "f32": (_abi.F32, np.float32, np.uint32),
           "bf16": (_abi.BF16, np.uint16, np.uint16)}


def operators(shape, kh, rng, uniform, upwind):
    jsize = tuple(reversed(shape))
    ops = []
    for N in range(1, 4):
        n = jsize[N - 1]
        h = 1.0 / (n + 1)
        dx = h if uniform in (True, "mixed") else h * (1 + 0.5 * rng.random(n + 1))
        ops.append(D.CenteredDifference[N](2, 2 * kh, dx, n))
        if upwind or (uniform == "mixed" and N == 1):
            ops.append(D.UpwindDifference[N](1, 3, dx, n, rng.standard_normal(n)))
    return ops, jsize


def values(shape, storage, rng):
    u = rng.uniform(-1, 1, shape)
    if storage == "bf16":
        return (u.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    return u.astype(STORAGE[storage][1])


def check(hc, d, u, storage, accumulate):
    _, dt, bits = STORAGE[storage]
    base = values(u.shape, storage, np.random.default_rng(5)) if accumulate else np.zeros_like(u)
    ref = base.copy()
    assert hc.hc_apply(C.byref(d), ref.ctypes.data, u.ctypes.data, accumulate) == 0, hc.hc_last_error()
    # sentinel: a NaN pattern no evaluation of these inputs produces
    sentinel = {"f64": 0x7FF8DEADBEEF0001, "f32": 0x7FC0DEAD, "bf16": 0x7FC1}[storage]
    got = base.copy()
    gv = got.view(bits)
    mark = np.zeros(u.shape, dtype=bool)
    if accumulate:
        # the accumulated value's sentinel check: run once on the sentinel-filled copy to find
        # the points written, once on the real base to compare
        probe = np.full(u.shape, sentinel, dtype=bits).view(got.dtype)
        nshell = hc.hc_shell(C.byref(d), probe.ctypes.data, u.ctypes.data, 0)
        mark = probe.view(bits) != sentinel
        assert hc.hc_shell(C.byref(d), got.ctypes.data, u.ctypes.data, 1) == nshell
    else:
        gv[...] = sentinel
        base = got.copy()
        nshell = hc.hc_shell(C.byref(d), got.ctypes.data, u.ctypes.data, 0)
        mark = gv != sentinel
    assert nshell > 0, "no 3-D tiled path for this plan"
    assert int(mark.sum()) == nshell
    # the shell is the complement of a box: every face plane of the grid is in it
    assert mark[0].all() and mark[-1].all() and mark[:, 0].all() and mark[:, -1].all()
    assert mark[:, :, 0].all() and mark[:, :, -1].all()
    np.testing.assert_array_equal(got.view(bits)[mark], ref.view(bits)[mark])
    # points the strips do not cover are left as they were
    np.testing.assert_array_equal(got.view(bits)[~mark], base.view(bits)[~mark])


CASES = [((32, 32, 32), 1), ((40, 48, 64), 2), ((33, 29, 70), 3), ((17, 100, 36), 4),
         ((12, 40, 600), 3), ((19, 23, 532), 2), ((24, 20, 64), 4), ((40, 30, 12), 2),
         ((20, 18, 16), 3)]


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("shape,kh", CASES)
@pytest.mark.parametrize("uniform", [True, False, "mixed"])
def test_shell_strips_equal_generic(hc, shape, kh, storage, uniform):
    rng = np.random.default_rng(100 + kh)
    ops, jsize = operators(shape, kh, rng, uniform, upwind=not uniform and kh == 3)
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.2), tuple(1.0 / (s + 1) for s in jsize),
                             1, jsize))
    bcs = _bcs_of(Q, 3, jsize)[0]
    d, keep = make_desc(ops, jsize, STORAGE[storage][0], input_kind=_abi.INPUT_GHOST,
                        semantics=_abi.SEM_LINES, bcs=bcs)
    u = values(shape, storage, rng)
    check(hc, d, u, storage, 0)
    check(hc, d, u, storage, 1)
    pts = kinds(hc, d)
    if storage != "f64" and shape[2] % 4 == 0 and shape[2] >= 32:
        # aligned rows: the faces take the bulk of the shell, the general items the edges
        assert pts[0] > 0 and pts[1] > 0 and pts[2] > 0, pts
        assert pts[3] < pts[:3].sum(), pts


@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("shape", [(20, 24, 64), (17, 33, 36)])
def test_shell_strips_wide_accumulation(hc, storage, shape):
    # third derivatives: DOP_ACC_AUTO sums f32 values in f64 (DOP_ACC_F32_WIDE), which the
    # tiled path and its shell now run too
    rng = np.random.default_rng(11)
    jsize = tuple(reversed(shape))
    ops = []
    for N in range(1, 4):
        n = jsize[N - 1]
        dx = (1.0 / (n + 1)) * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](3, 4, dx, n), D.CenteredDifference[N](2, 2, dx, n)]
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.2), tuple(1.0 / (s + 1) for s in jsize),
                             1, jsize))
    bcs = _bcs_of(Q, 3, jsize)[0]
    d, keep = make_desc(ops, jsize, STORAGE[storage][0], input_kind=_abi.INPUT_GHOST,
                        semantics=_abi.SEM_LINES, bcs=bcs)
    assert d.accum == _abi.ACC_AUTO
    u = values(shape, storage, rng)
    check(hc, d, u, storage, 0)
    check(hc, d, u, storage, 1)


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("periodic", [0.0, 0.3])
def test_shell_strips_per_line_bcs(hc, storage, periodic):
    # one atomic BC per line along x (periodic lines take no staged line ends)
    rng = np.random.default_rng(7)
    shape = (20, 18, 64)
    ops, jsize = operators(shape, 2, rng, False, upwind=False)
    arr = np.empty((jsize[1], jsize[2]), dtype=object)
    for i in np.ndindex(*arr.shape):
        r = rng.random()
        arr[i] = (D.PeriodicBC(float) if r < periodic else
                  D.RobinBC(tuple(rng.uniform(0.5, 1.5, 3)), tuple(rng.uniform(0.5, 1.5, 3)), 0.02, 1)
                  if r < 0.7 else D.DirichletBC(float(rng.standard_normal()), 0.5))
    Q = D.compose(D.MultiDimBC[1](arr), D.MultiDimBC[2](D.Dirichlet0BC(float), jsize),
                  D.MultiDimBC[3](D.NeumannBC((0.1, -0.3), 0.05, 1), jsize))
    bcs = _bcs_of(Q, 3, jsize)[0]
    d, keep = make_desc(ops, jsize, STORAGE[storage][0], input_kind=_abi.INPUT_GHOST,
                        semantics=_abi.SEM_LINES, bcs=bcs)
    u = values(shape, storage, rng)
    check(hc, d, u, storage, 0)
    check(hc, d, u, storage, 1)
