"""CPU tests of the product's host side: the Python mirror of the reference constructors
(tables must equal the oracle's bit for bit), the C ABI (library loads, exports every
declared symbol, struct layouts agree), and the device path's plan builder + per-point
arithmetic compiled for the CPU (tests/hostcheck) against the FP64 oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import make_desc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC_PATH = os.environ.get("DOP_HOSTCHECK_LIBRARY") or os.path.join(ROOT, "tests", "hostcheck", "_build",
                                                                "libdop_hostcheck.so")


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    return np.abs(np.asarray(a, dtype=np.float64) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC_PATH):
        pytest.skip("tests/hostcheck not built (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(HC_PATH)
    L.hc_apply.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p, C.c_void_p, C.c_int]
    L.hc_plan_info.argtypes = [C.POINTER(_abi.DopDesc), C.POINTER(C.c_int32)]
    L.hc_last_error.restype = C.c_char_p
    L.hc_sizeof.restype = C.c_int64
    return L


# ---------------------------------------------------------------------------------------
# tables of the Python constructors == oracle tables (bitwise)

CASES = []
for wind in (0, 1):
    for d in (1, 2, 3, 4):
        for a in (1, 2, 3, 4, 6, 8):
            if wind == 0 and a < 2:
                continue
            for off in ((0, 1, 2, 3) if wind else (0,)):
                if wind and off > (d + a - 1) // 2:
                    continue
                for uni in (True, False):
                    CASES.append((wind, d, a, off, uni))


@pytest.mark.parametrize("wind,d,a,off,uni", CASES)
def test_constructor_tables_match_oracle(wind, d, a, off, uni):
    rng = np.random.default_rng(d * 100 + a * 10 + off)
    n = 40
    dx = 0.1 if uni else 0.05 * (1 + 0.5 * rng.random(n + 1))
    o = po.Op(wind, d, a, dx, n, offside=off)
    P = D.UpwindDifference(d, a, dx, n, offside=off) if wind else D.CenteredDifference(d, a, dx, n)
    assert (P.stencil_length, P.boundary_stencil_length, P.boundary_point_count) == (o.SL, o.BSL, o.BPC)
    for x, y in ((o.stencil, P.stencil_coefs), (o.low, P.low_boundary_coefs), (o.high, P.high_boundary_coefs)):
        y = y.ravel()
        assert np.array_equal(x.ravel()[: y.size], y)


def test_robin_tables_match_oracle(rng):
    for order in (1, 2, 3):
        for _ in range(4):
            l = tuple(rng.random(3) + 0.1)
            r = tuple(rng.random(3) + 0.1)
            dx = float(rng.random() + 0.01)
            for dxa in (dx, dx * (1 + rng.random(20))):
                a = D.RobinBC(l, r, dxa, order)
                b = po.robin(l, r, dxa, order)
                assert np.array_equal(a.a_l, b.a_l) and np.array_equal(a.a_r, b.a_r)
                assert a.b_l == b.b_l and a.b_r == b.b_r
    a, b = D.Dirichlet0BC(float), po.dirichlet0()
    assert np.array_equal(a.a_l, b.a_l) and a.b_l == b.b_l


def test_constructor_errors():
    with pytest.raises(AssertionError):
        D.CenteredDifference(2, 1, 0.1, 10)          # derivative_operator.jl:84
    with pytest.raises(AssertionError):
        D.UpwindDifference(1, 2, 0.1, 10, offside=2)  # :459
    with pytest.raises(ValueError):
        (2.0 * np.ones(3)) * D.CenteredDifference(2, 2, 0.1, 10)   # DimensionMismatch
    Qs = D.Dirichlet0BC(float, (8, 9))
    with pytest.raises(ValueError):
        D.compose(Qs[0], Qs[0])
    with pytest.raises(ValueError):
        D.compose(Qs[0])
    with pytest.raises(ValueError):
        D.GeneralBC([1.0], [1.0, 2.0], 0.1, 2)


def test_scaling_keeps_coeff_func():
    L = D.UpwindDifference(1, 2, 0.1 * np.ones(11), 10, 1.0)
    M = -1.0 * L
    assert np.all(M.coefficients == -1.0)
    assert M.coeff_kind == _abi.COEFF_NUM_POS   # coeff_func stays 1.0 (quirk of c*A)
    V = np.linspace(-1, 1, 10) * L
    assert np.allclose(V.coefficients, np.linspace(-1, 1, 10))


def test_tensor_api_has_no_cpu_path():
    import torch
    L = D.CenteredDifference(2, 2, 0.1, 10)
    with pytest.raises(TypeError):
        (L * D.Dirichlet0BC(float)) * torch.zeros(10, dtype=torch.float64)


# ---------------------------------------------------------------------------------------
# the C ABI

def _declared_functions():
    text = open(os.path.join(ROOT, "include", "dop.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|int64_t|double|const char\*)\s+\**(dop_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert "dop_apply" in names and "dop_plan_create" in names and len(names) >= 12
    lib = C.CDLL(D.library_path())
    for n in names:
        assert hasattr(lib, n), n
    assert lib.dop_version() == _abi.ABI_VERSION


def test_struct_layouts_match(hc):
    assert hc.hc_sizeof(0) == C.sizeof(_abi.DopOperator)
    assert hc.hc_sizeof(1) == C.sizeof(_abi.DopAffineBC)
    assert hc.hc_sizeof(2) == C.sizeof(_abi.DopDesc)


# ---------------------------------------------------------------------------------------
# device arithmetic compiled for the CPU vs the oracle

def hc_apply(hc, terms, u, *, storage=_abi.F64, input_kind=_abi.INPUT_GHOST, sem=_abi.SEM_LINES,
             bcs=None, padded=(0, 0, 0), tv=0, out_shape=None, accum=_abi.ACC_AUTO, du=None):
    jsize = tuple(reversed(out_shape if out_shape is not None else u.shape))
    d, keep = make_desc(terms, jsize, storage, input_kind=input_kind, semantics=sem, bcs=bcs,
                        padded=padded, transverse_view=tv, accum=accum)
    acc = 0
    if du is None:
        du = np.zeros(out_shape if out_shape is not None else u.shape, dtype=u.dtype)
    else:
        acc = 1
    rc = hc.hc_apply(C.byref(d), du.ctypes.data, u.ctypes.data, acc)
    if rc != 0:
        raise ValueError(f"{rc}: {hc.hc_last_error().decode()}")
    return du


def _oracle_op(L, axis):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=axis, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


def _obc(bc):
    from diffeqoperators_amd.apply import _PeriodicPad
    if isinstance(bc, _PeriodicPad):
        return po.periodic() if bc.kind == _abi.BC_PERIODIC else po.periodic_nd()
    return po.BC(bc.a_l, bc.b_l, bc.a_r, bc.b_r)


@pytest.mark.parametrize("d,a", [(1, 2), (2, 2), (2, 4), (3, 2), (4, 4), (2, 6), (4, 8), (1, 8)])
@pytest.mark.parametrize("coeff", ["const", "vector"])
def test_hc_1d_centered_bpv(hc, rng, d, a, coeff):
    n = 37
    u = rng.standard_normal(n)
    c = 1.0 if coeff == "const" else rng.standard_normal(n)
    L = D.CenteredDifference(d, a, 0.1, n, c)
    bc = D.RobinBC((1.0, 0.5, 0.3), (0.7, 0.2, -0.1), 0.1, 2)
    got = hc_apply(hc, [L], u, sem=_abi.SEM_BPV, bcs={0: bc})
    ref = po.ghost_apply([_oracle_op(L, 0)], [_obc(bc)], u, bpv_1d=True)
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("d,a,off", [(1, 1, 0), (1, 2, 0), (1, 2, 1), (2, 2, 0), (2, 3, 0),
                                     (2, 3, 1), (2, 3, 2), (1, 3, 0), (3, 2, 0), (1, 4, 2)])
@pytest.mark.parametrize("grid", ["uniform", "irregular"])
@pytest.mark.parametrize("wind", ["pos", "neg", "mixed", "scaled"])
@pytest.mark.parametrize("sem", [_abi.SEM_BPV, _abi.SEM_LINES])
def test_hc_1d_upwind(hc, rng, d, a, off, grid, wind, sem):
    n = 23
    dx = 0.1 if grid == "uniform" else 0.05 * (1 + 0.5 * rng.random(n + 1))
    c = {"pos": 1.3, "neg": -0.7, "mixed": rng.standard_normal(n), "scaled": 1.0}[wind]
    L = D.UpwindDifference(d, a, dx, n, c, offside=off)
    if wind == "scaled":
        L = -2.0 * L     # coefficients negative, coeff_func still positive
    u = rng.standard_normal(n)
    bc = D.NeumannBC((0.3, -0.2), dx, 1)
    got = hc_apply(hc, [L], u, sem=sem, bcs={0: bc})
    ref = po.ghost_apply([_oracle_op(L, 0)], [_obc(bc)], u, bpv_1d=(sem == _abi.SEM_BPV))
    assert rel(got, ref) < 1e-13


def test_hc_1d_combination_same_axis(hc, rng):
    n = 50
    u = rng.standard_normal(n)
    L1 = D.CenteredDifference(2, 4, 0.02, n)
    L2 = D.UpwindDifference(1, 3, 0.02, n, rng.standard_normal(n))
    bc = D.Dirichlet0BC(float)
    got = hc_apply(hc, [L1, L2], u, sem=_abi.SEM_BPV, bcs={0: bc})
    ref = po.ghost_apply([_oracle_op(L1, 0), _oracle_op(L2, 0)], [_obc(bc)], u)
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("shape", [(29, 31), (8, 70), (11, 9, 13), (6, 7, 40)])
def test_hc_nd_ghost(hc, rng, shape):
    u = rng.standard_normal(shape)
    jsize = tuple(reversed(shape))
    nd = len(shape)
    terms, oterms = [], []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        Lc = D.CenteredDifference[N](2, 6 if n >= 8 else 2, dx, n)
        Lu = D.UpwindDifference[N](1, 3, dx, n, -rng.standard_normal(n))
        terms += [Lc, Lu]
        oterms += [_oracle_op(Lc, N - 1), _oracle_op(Lu, N - 1)]
    bcs = {0: D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), 0.05)}
    for k in range(1, nd):
        bcs[k] = D.Dirichlet0BC(float)
    got = hc_apply(hc, terms, u, bcs=bcs)
    ref = po.ghost_apply(oterms, [_obc(bcs[k]) for k in range(nd)], u)
    assert rel(got, ref) < 1e-13


def test_hc_accumulate(hc, rng):
    u = rng.standard_normal((20, 30))
    du0 = rng.standard_normal((20, 30))
    L = D.CenteredDifference[2](2, 4, 0.1, 20)
    bc = D.Dirichlet0BC(float)
    got = hc_apply(hc, [L], u, bcs={1: bc}, du=du0.copy())
    ref = po.ghost_apply([_oracle_op(L, 1)], [_obc(bc), _obc(bc)], u) + du0
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("pad_other", [0, 1])
def test_hc_padded_single_operator(hc, rng, pad_other):
    # mul!(du, L, M) with M padded along L's axis and (optionally) trimmed elsewhere
    n1, n2 = 24, 17
    L = D.CenteredDifference[1](4, 4, 0.1, n1, rng.standard_normal(n1))
    M = rng.standard_normal((n2 + 2 * pad_other, n1 + 2))
    got = hc_apply(hc, [L], M, input_kind=_abi.INPUT_PADDED, padded=(1, pad_other, 0),
                   out_shape=(n2, n1))
    ref = po.mul_nd(_oracle_op(L, 0), M, (n2, n1))
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("nd", [2, 3])
def test_hc_composite_transverse_view(hc, rng, nd):
    n = 10
    M = rng.standard_normal((n + 2,) * nd)
    ops = [D.CenteredDifference[k](2, 2, 1.0, n) for k in range(1, nd + 1)]
    ops.append(D.UpwindDifference[1](1, 2, 1.0, n, rng.standard_normal(n)))
    got = hc_apply(hc, ops, M, input_kind=_abi.INPUT_PADDED, padded=(1,) * nd + (0,) * (3 - nd),
                   tv=1, out_shape=(n,) * nd)
    ref = po.composite_mul_padded([_oracle_op(L, L.N - 1) for L in ops], M)
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("storage,tol", [(_abi.F32, 1e-6), (_abi.BF16, 1e-2)])
def test_hc_reduced_storage(hc, rng, storage, tol):
    import torch
    shape = (12, 14, 16)
    u64 = rng.uniform(-1, 1, shape)
    tdt = torch.float32 if storage == _abi.F32 else torch.bfloat16
    uq = torch.from_numpy(u64).to(tdt)
    uq64 = uq.to(torch.float64).numpy()       # the quantized input, exactly
    terms, oterms = [], []
    for N in range(1, 4):
        n = shape[3 - N]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        Lc = D.CenteredDifference[N](2, 6, dx, n)
        terms.append(Lc)
        oterms.append(_oracle_op(Lc, N - 1))
    bc = D.Dirichlet0BC(float)
    if storage == _abi.F32:
        u_in = uq.numpy()
        got = hc_apply(hc, terms, u_in, storage=storage, bcs={0: bc, 1: bc, 2: bc}).astype(np.float64)
    else:
        bits = uq.view(torch.int16).numpy().copy()
        out = hc_apply(hc, terms, bits, storage=storage, bcs={0: bc, 1: bc, 2: bc})
        got = torch.from_numpy(out).view(torch.bfloat16).to(torch.float64).numpy()
    ref = po.ghost_apply(oterms, [_obc(bc)] * 3, uq64)
    assert rel(got, ref) < tol


def test_hc_kahan_high_order(hc, rng):
    # d = 4, FP32 storage: compensated accumulation keeps the sum close to the FP64 result
    n = 64
    x = np.linspace(0, 1, n)
    u = np.sin(3 * x).astype(np.float32)
    L = D.CenteredDifference(4, 8, 1.0 / (n + 1), n)
    bc = D.Dirichlet0BC(float)
    ref = po.ghost_apply([_oracle_op(L, 0)], [_obc(bc)], u.astype(np.float64))
    k = hc_apply(hc, [L], u, storage=_abi.F32, sem=_abi.SEM_BPV, bcs={0: bc}, accum=_abi.ACC_F32_KAHAN)
    p = hc_apply(hc, [L], u, storage=_abi.F32, sem=_abi.SEM_BPV, bcs={0: bc}, accum=_abi.ACC_F32)
    assert rel(k, ref) < 1e-5
    assert rel(k, ref) <= rel(p, ref)
    # exact f32 products summed in f64: at least as close, and what the automatic mode
    # picks for d >= 3
    w = hc_apply(hc, [L], u, storage=_abi.F32, sem=_abi.SEM_BPV, bcs={0: bc}, accum=_abi.ACC_F32_WIDE)
    a = hc_apply(hc, [L], u, storage=_abi.F32, sem=_abi.SEM_BPV, bcs={0: bc}, accum=_abi.ACC_AUTO)
    assert rel(w, ref) < 1e-5
    assert rel(w, ref) <= rel(p, ref)
    assert np.array_equal(w, a)


def test_hc_plan_rejects_bad_shapes(hc, rng):
    L = D.CenteredDifference(2, 6, 0.1, 5)   # BPC = 2: len 5 is too short for the reference
    with pytest.raises(ValueError):
        hc_apply(hc, [L], rng.standard_normal(5), sem=_abi.SEM_BPV, bcs={0: D.Dirichlet0BC(float)})
    L = D.CenteredDifference[2](2, 2, 0.1, 10)
    with pytest.raises(ValueError):   # no BC on axis 2
        hc_apply(hc, [L], rng.standard_normal((10, 10)), bcs={0: D.Dirichlet0BC(float)})


# ---------------------------------------------------------------------------------------
# PeriodicBC: which padding each form of Q selects, and the device arithmetic

def test_periodic_bc_forms():
    from diffeqoperators_amd.apply import _bcs_of
    P = D.PeriodicBC(float)
    bcs, sem = _bcs_of(P, 1, (10,))
    assert sem == _abi.SEM_BPV and bcs[0].kind == _abi.BC_PERIODIC
    bcs, sem = _bcs_of(P, 2, (10, 6))            # BC*u on a matrix: MultiDimBC{1}
    assert bcs[0].kind == _abi.BC_PERIODIC_ND
    Qx, Qy = D.PeriodicBC(float, (10, 6))
    bcs, _ = _bcs_of(D.compose(Qx, Qy), 2, (10, 6))
    assert bcs[0].kind == bcs[1].kind == _abi.BC_PERIODIC_ND
    bcs, _ = _bcs_of(D.PeriodicBC[2](float, (10, 6)), 2, (10, 6))
    assert bcs[1].kind == _abi.BC_PERIODIC_ND
    mixed = D.compose(D.PeriodicBC[1](float, (10, 6)), D.Dirichlet0BC[2](float, (10, 6)))
    bcs, _ = _bcs_of(mixed, 2, (10, 6))
    assert bcs[0].kind == _abi.BC_PERIODIC and isinstance(bcs[1], D.AffineBC)


@pytest.mark.parametrize("wind,d,a", [(0, 2, 2), (0, 1, 4), (0, 4, 6), (1, 1, 2), (1, 2, 3)])
def test_hc_1d_periodic_bpv(hc, rng, wind, d, a):
    from diffeqoperators_amd.apply import _bcs_of
    n = 31
    u = rng.standard_normal(n)
    L = D.UpwindDifference(d, a, 0.1, n, -1.0) if wind else D.CenteredDifference(d, a, 0.1, n)
    bcs, sem = _bcs_of(D.PeriodicBC(float), 1, (n,))
    got = hc_apply(hc, [L], u, sem=sem, bcs=bcs)
    ref = po.ghost_apply([_oracle_op(L, 0)], [po.periodic()], u, bpv_1d=True)
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("form", ["all", "mixed"])
@pytest.mark.parametrize("shape", [(17, 23), (9, 10, 14)])
def test_hc_nd_periodic(hc, rng, form, shape):
    from diffeqoperators_amd.apply import _bcs_of
    u = rng.standard_normal(shape)
    js = tuple(reversed(shape))
    nd = len(shape)
    if form == "all":
        Q = D.compose(*D.PeriodicBC(float, js))
    else:
        Q = D.compose(D.PeriodicBC[1](float, js), *[D.Dirichlet0BC[k](float, js) for k in range(2, nd + 1)])
    bcs, sem = _bcs_of(Q, nd, js)
    terms = [D.CenteredDifference[N](2, 4, 0.1, js[N - 1]) for N in range(1, nd + 1)]
    got = hc_apply(hc, terms, u, sem=sem, bcs=bcs)
    ref = po.ghost_apply([_oracle_op(L, L.N - 1) for L in terms], [_obc(bcs[k]) for k in range(nd)], u)
    assert rel(got, ref) < 1e-13


# ---------------------------------------------------------------------------------------
# GeneralBC: tables equal the oracle's bit for bit; applied through the device arithmetic

GENERAL_CASES = [([-10.0, 1.0, 6.0], [-10.0, 1.0, 6.0], 1.0, 3),
                 ([0.0, 1.0, -0.3, 0.0, -1.0], [0.0, 1.0, -0.2, 0.0, -1.0], 0.05, 3),   # KdV.jl:28
                 ([0.5, 1.0, 2.0, 3.0], [0.2, 1.0, -1.0, 0.5], "vector", 2)]


@pytest.mark.parametrize("al,ar,dx,order", GENERAL_CASES)
def test_general_bc_tables_match_oracle(al, ar, dx, order):
    if dx == "vector":
        dx = np.linspace(0.1, 0.2, 20)
    g = D.GeneralBC(al, ar, dx, order)
    q = po.general(al, ar, dx, order)
    assert np.array_equal(g.a_l, q.a_l) and np.array_equal(g.a_r, q.a_r)
    assert g.b_l == q.b_l and g.b_r == q.b_r


@pytest.mark.parametrize("al,ar,dx,order", GENERAL_CASES)
def test_hc_general_bc(hc, rng, al, ar, dx, order):
    n = 30
    dxv = np.linspace(0.1, 0.2, n + 1) if dx == "vector" else dx
    bc = D.GeneralBC(al, ar, dxv, order)
    u = rng.standard_normal(n)
    L = D.CenteredDifference(3, 2, dxv, n) if dx != "vector" else D.CenteredDifference(2, 2, dxv, n)
    got = hc_apply(hc, [L], u, sem=_abi.SEM_BPV, bcs={0: bc})
    ref = po.ghost_apply([_oracle_op(L, 0)], [_obc(bc)], u, bpv_1d=True)
    assert rel(got, ref) < 1e-13
    # 2-D: the same atom extended over both axes
    U = rng.standard_normal((n, n))
    ops = [D.CenteredDifference[1](2, 2, dxv, n), D.CenteredDifference[2](2, 2, dxv, n)]
    got = hc_apply(hc, ops, U, bcs={0: bc, 1: bc})
    ref = po.ghost_apply([_oracle_op(L2, L2.N - 1) for L2 in ops], [_obc(bc), _obc(bc)], U)
    assert rel(got, ref) < 1e-13


# ---------------------------------------------------------------------------------------
# buffers handed to the batched and RK4 entry points are checked before anything is created

def test_batch_and_rk4_reject_host_tensors():
    h = 1.0 / 33
    A = D.CenteredDifference(2, 2, h, 32) * D.Dirichlet0BC(float)
    u = torch.zeros(32, dtype=torch.float64)
    with pytest.raises(TypeError):
        D.Batch([A], [u], [torch.zeros_like(u)])
    with pytest.raises(TypeError):
        D.RK4([A], [u], [1.0], [1e-4])
    with pytest.raises(ValueError):
        D.Batch([A], [u], [])
