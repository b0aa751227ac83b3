"""Per-line boundary conditions, MultiDimBC{dim}(BC::Array) (multi_dim_bc_operators.jl:93-95,
applied line by line by slice_rmul, :34): CPU tests.  The oracle's per-line padding is pinned by
an explicit loop over the lines (each line's 1-D Q*u and operator, numpy slicing: this fixes
the line order); the device path's arithmetic compiled for the CPU (tests/hostcheck) and its
concretization are checked against it; BoundaryPadded.materialize against the ghost formula."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import _PeriodicPad, _bcs_of, make_desc
from diffeqoperators_amd.bcs import PeriodicAtomicBC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC_PATH = os.environ.get("DOP_HOSTCHECK_LIBRARY") or os.path.join(ROOT, "tests", "hostcheck", "_build",
                                                                "libdop_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC_PATH):
        pytest.skip("tests/hostcheck not built")
    L = C.CDLL(HC_PATH)
    L.hc_apply.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p, C.c_void_p, C.c_int]
    L.hc_csr.argtypes = [C.POINTER(_abi.DopDesc)] + [C.c_void_p] * 4 + [C.POINTER(C.c_int64)]
    L.hc_last_error.restype = C.c_char_p
    return L


def random_bc(rng, dx):
    k = rng.integers(0, 4)
    if k == 0:
        return D.RobinBC(tuple(rng.uniform(0.5, 1.5, 3)), tuple(rng.uniform(0.5, 1.5, 3)), dx,
                         int(rng.integers(1, 4)))
    if k == 1:
        return D.DirichletBC(float(rng.standard_normal()), float(rng.standard_normal()))
    if k == 2:
        return D.NeumannBC((float(rng.standard_normal()), float(rng.standard_normal())), dx, 2)
    return D.GeneralBC([rng.uniform(0.5, 1.0), 1.0, 0.3], [rng.uniform(0.5, 1.0), 1.0, -0.2], dx, 2)


def bc_array(rng, shape, dx, periodic=0.0):
    """Random atomic BCs, a fraction `periodic` of them PeriodicBC."""
    a = np.empty(shape, dtype=object)
    for i in np.ndindex(*shape):
        a[i] = D.PeriodicBC(float) if rng.random() < periodic else random_bc(rng, dx)
    return a


def obc(b):
    if isinstance(b, PeriodicAtomicBC):
        return po.periodic()   # the line wraps (bc_operators.jl:192)
    return po.BC(b.a_l, b.b_l, b.a_r, b.b_r)


def oop(L):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=L.N - 1, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def test_oracle_per_line_equals_line_loop(rng):
    # 3-D, per-line BCs along Julia axis 2: lines indexed by (i1, i3)
    n1, n2, n3 = 5, 9, 4
    u = rng.standard_normal((n3, n2, n1))
    L = D.CenteredDifference[2](2, 4, 0.1, n2, rng.uniform(0.5, 1.5, n2))
    arr = bc_array(rng, (n1, n3), 0.1)                      # Julia order of the other axes
    lines = [obc(b) for b in arr.flatten(order="F")]       # column-major: i1 fastest
    got = po.ghost_apply([oop(L)], [po.dirichlet0(), po.BC.per_line(lines), po.dirichlet0()], u,
                         bpv_1d=False)
    ref = np.zeros_like(u)
    for i3 in range(n3):
        for i1 in range(n1):
            line = u[i3, :, i1]
            ref[i3, :, i1] = po.ghost_apply([oop(D.CenteredDifference(2, 4, 0.1, n2, L.coefficients))],
                                            [obc(arr[i1, i3])], line, bpv_1d=False)
    assert np.array_equal(got, ref)


def test_oracle_per_line_periodic_pattern(rng):
    # multi_dim_bc_test.jl:7-35: half the lines Robin, half PeriodicBC; each padded line equals
    # its own atomic BC applied to that line (Ax[:, j] == Array(BCx[j] * A[:, j]))
    n, m = 8, 15
    A = rng.random((m, n))                                  # Julia size (n, m)
    q1 = D.RobinBC((1.0, 2.0, 3.0), (0.0, -1.0, 2.0), 0.1, 4)
    q2 = D.PeriodicBC(float)
    BCx = [q1] * (m // 2) + [q2] * (m - m // 2)
    L = D.CenteredDifference[1](2, 2, 0.1, n)
    got = po.ghost_apply([oop(L)], [po.BC.per_line([obc(b) for b in BCx]), po.dirichlet0()], A,
                         bpv_1d=False)
    for j in range(m):
        line = A[j, :]
        gl, gr = (line[-1], line[0]) if j >= m // 2 else obc(q1).pad(line)
        padded = np.concatenate([[gl], line, [gr]])
        ref = oop(L).mul_padded(padded)
        # the ghost of a Robin line is summed in another order here (numpy's dot)
        assert np.abs(got[j, :] - ref).max() <= 1e-13 * np.abs(ref).max(), j


def test_all_periodic_lines_take_the_nd_padding():
    # an array whose eltype is PeriodicBC takes the N-D method (multi_dim_bc_operators.jl:221-228)
    arr = np.empty((6,), dtype=object)
    for i in range(6):
        arr[i] = D.PeriodicBC(float)
    bcs, _ = _bcs_of(D.MultiDimBC[1](arr), 2, (5, 6))
    assert isinstance(bcs[0], _PeriodicPad) and bcs[0].kind == _abi.BC_PERIODIC_ND
    # mixed: one BC per line
    arr[2] = D.Dirichlet0BC(float)
    bcs, _ = _bcs_of(D.MultiDimBC[1](arr), 2, (5, 6))
    assert not isinstance(bcs[0], _PeriodicPad)


@pytest.mark.parametrize("shape,periodic", [((12, 17), 0.0), ((5, 6, 7), 0.0), ((12, 17), 0.5),
                                             ((5, 6, 7), 0.4)])
def test_hostcheck_per_line_matches_oracle(hc, rng, shape, periodic):
    jsize = tuple(reversed(shape))
    nd = len(shape)
    u = rng.standard_normal(shape)
    ops, oops, bcs, obcs = [], [], {}, []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](2, 4, dx, n), D.UpwindDifference[N](1, 2, dx, n, rng.standard_normal(n))]
        oops += [oop(ops[-2]), oop(ops[-1])]
        other = tuple(jsize[k] for k in range(nd) if k != N - 1)
        arr = bc_array(rng, other, 0.05, periodic)
        Q = D.MultiDimBC[N](arr)
        bcs.update(_bcs_of(Q, nd, jsize)[0])
        obcs.append(po.BC.per_line([obc(b) for b in arr.flatten(order="F")]))
    d, keep = make_desc(ops, jsize, _abi.F64, input_kind=_abi.INPUT_GHOST,
                        semantics=_abi.SEM_LINES, bcs=bcs)
    du = np.zeros_like(u)
    assert hc.hc_apply(C.byref(d), du.ctypes.data, u.ctypes.data, 0) == 0, hc.hc_last_error()
    ref = po.ghost_apply(oops, obcs, u)
    assert rel(du, ref) < 1e-13
    # concretization: columns against the oracle's apply
    n = u.size
    rp = np.zeros(n + 1, dtype=np.int64)
    nnz = C.c_int64()
    assert hc.hc_csr(C.byref(d), rp.ctypes.data, None, None, None, C.byref(nnz)) == 0
    col = np.zeros(nnz.value, dtype=np.int64)
    val = np.zeros(nnz.value)
    b = np.zeros(n)
    assert hc.hc_csr(C.byref(d), rp.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data,
                     C.byref(nnz)) == 0
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    z = po.ghost_apply(oops, obcs, np.zeros(shape)).ravel()
    assert np.abs(b - z).max() <= 1e-12 * max(np.abs(z).max(), 1.0)
    for j in range(0, n, max(1, n // 40)):
        e = np.zeros(n)
        e[j] = 1.0
        cj = po.ghost_apply(oops, obcs, e.reshape(shape)).ravel() - z
        assert np.abs(A[:, j] - cj).max() <= 1e-12 * np.abs(A).max()


def test_materialize_per_line(rng):
    n1, n2 = 7, 5
    u = torch.from_numpy(rng.standard_normal((n2, n1)))
    arr = bc_array(rng, (n2,), 0.1)           # BCs along Julia axis 1: one per i2
    Q = D.compose(D.MultiDimBC[1](arr), D.MultiDimBC[2](D.Dirichlet0BC(float), (n1, n2)))
    P = (Q * u).materialize().numpy()
    assert P.shape == (n2 + 2, n1 + 2)
    un = u.numpy()
    for i2 in range(n2):
        b = arr[i2]
        gl = np.dot(b.a_l, un[i2, : len(b.a_l)]) + b.b_l
        gr = np.dot(b.a_r, un[i2, n1 - len(b.a_r):]) + b.b_r
        assert P[i2 + 1, 0] == pytest.approx(gl, rel=1e-14, abs=1e-14)
        assert P[i2 + 1, -1] == pytest.approx(gr, rel=1e-14, abs=1e-14)


def test_materialize_per_line_periodic(rng):
    n1, n2 = 7, 6
    u = torch.from_numpy(rng.standard_normal((n2, n1)))
    arr = bc_array(rng, (n2,), 0.1, periodic=0.5)
    arr[0], arr[1] = D.PeriodicBC(float), D.Dirichlet0BC(float)
    P = (D.MultiDimBC[1](arr) * u).materialize().numpy()
    un = u.numpy()
    for i2 in range(n2):
        gl, gr = obc(arr[i2]).pad(un[i2])
        assert P[i2, 0] == pytest.approx(gl, rel=1e-14, abs=1e-14)
        assert P[i2, -1] == pytest.approx(gr, rel=1e-14, abs=1e-14)


def test_per_line_size_mismatch_raises():
    arr = np.empty((4,), dtype=object)
    for i in range(4):
        arr[i] = D.Dirichlet0BC(float)
    Q = D.compose(D.MultiDimBC[1](arr), D.MultiDimBC[2](D.Dirichlet0BC(float), (6, 5)))
    with pytest.raises(AssertionError):
        _bcs_of(Q, 2, (6, 5))
