"""CPU tests of the concretization (the matrix form of a plan) exported by libdop: the same
row enumeration and merge (csrc/dop_concrete.hpp) compiled for the CPU (tests/hostcheck), checked
against the reference tests' golden matrices (`Array(L) == correct`,
derivative_operators_interface.jl:121-143, :229-281; upwind_operators_interface.jl:152-839),
against the reference's own composition formulas (L * Q_L, L * Q_b,
concretization.jl:866-870, with Q_L / Q_b of concretization.jl:174-180, :210-213; the Kron form
of N-D operators, :425-444), and column by column against the oracle's apply on unit vectors."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import _PeriodicPad, make_desc

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
    L.hc_csr.argtypes = [C.POINTER(_abi.DopDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.POINTER(C.c_int64)]
    L.hc_last_error.restype = C.c_char_p
    return L


def hc_csr(hc, d, n_rows):
    """(row_ptr, col, val, b) of the plan described by d."""
    rp = np.zeros(n_rows + 1, dtype=np.int64)
    nnz = C.c_int64()
    rc = hc.hc_csr(C.byref(d), rp.ctypes.data, None, None, None, C.byref(nnz))
    assert rc == 0, hc.hc_last_error().decode()
    col = np.zeros(max(nnz.value, 1), dtype=np.int64)
    val = np.zeros(max(nnz.value, 1), dtype=np.float64)
    b = np.zeros(n_rows)
    rc = hc.hc_csr(C.byref(d), rp.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data,
                   C.byref(nnz))
    assert rc == 0
    return rp, col[: nnz.value], val[: nnz.value], b


def dense(rp, col, val, n_rows, n_cols):
    A = np.zeros((n_rows, n_cols))
    for i in range(n_rows):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


def check_structure(rp, col, val):
    assert np.all(np.diff(rp) >= 0)
    for i in range(len(rp) - 1):
        c = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0)          # ascending, merged
    assert np.all(val != 0.0)                   # exact zeros are not stored


def padded_matrix(hc, ops, msz):
    """Matrix of the plain `L*M` over M of Julia size msz (padded along the ops' axis)."""
    N = ops[0].N
    osz = list(msz)
    osz[N - 1] -= 2
    padded = [0, 0, 0]
    padded[N - 1] = 1
    d, keep = make_desc(ops, tuple(osz), _abi.F64, input_kind=_abi.INPUT_PADDED,
                        semantics=_abi.SEM_LINES, padded=tuple(padded))
    n_rows, n_cols = int(np.prod(osz)), int(np.prod(msz))
    rp, col, val, b = hc_csr(hc, d, n_rows)
    check_structure(rp, col, val)
    assert np.all(b == 0.0)
    return dense(rp, col, val, n_rows, n_cols)


def ghost_matrix(hc, ops, bcs, jsize, sem):
    d, keep = make_desc(ops, tuple(jsize), _abi.F64, input_kind=_abi.INPUT_GHOST, semantics=sem,
                        bcs=bcs)
    n = int(np.prod(jsize))
    rp, col, val, b = hc_csr(hc, d, n)
    check_structure(rp, col, val)
    return dense(rp, col, val, n, n), b


# ---------------------------------------------------------------------------------------
# golden matrices: Array(L) == correct in the reference tests

def test_centered_uniform_golden(hc, fixtures):
    for e in fixtures["centered_uniform"]:
        L = D.CenteredDifference(e["d"], e["a"], e["dx"], e["N"])
        assert rel(padded_matrix(hc, [L], (e["N"] + 2,)), np.array(e["A"])) < 1e-14, e["name"]


def test_centered_irregular_golden(hc, fixtures):
    irr = fixtures["centered_irregular"]
    for e in irr["entries"]:
        n = len(irr["x"]) - 2
        L = D.CenteredDifference(e["d"], e["a"], np.array(irr["dx"]), n)
        M = padded_matrix(hc, [L], (n + 2,))
        if e["rows"] == "interior":
            M = M[1:-1]
        assert rel(M, np.array(e["A"])) < 1e-13, e["name"]


@pytest.mark.parametrize("group", ["upwind_uniform", "upwind_irregular"])
def test_upwind_golden(hc, fixtures, group):
    for e in fixtures[group]:
        n = len(e["A"])
        dx = e["dx"] if np.ndim(e["dx"]) == 0 else np.array(e["dx"])
        coeff = e["coeff"] if np.ndim(e["coeff"]) == 0 else np.array(e["coeff"])
        L = D.UpwindDifference(e["d"], e["a"], dx, n, coeff)
        assert rel(padded_matrix(hc, [L], (n + 2,)), np.array(e["A"])) < 1e-13, e["name"]


def test_planted_mistake_is_caught(hc, fixtures):
    # a wrong approximation order must miss the golden matrix
    e = fixtures["centered_uniform"][0]
    L = D.CenteredDifference(e["d"], e["a"] + 2, e["dx"], e["N"])
    assert rel(padded_matrix(hc, [L], (e["N"] + 2,)), np.array(e["A"])) > 1e-3


# ---------------------------------------------------------------------------------------
# the reference's composition formulas

def q_matrices(bc, n):
    """(Q_L, Q_b) of an atomic BC on a line of n (concretization.jl:174-180, :210-213)."""
    QL = np.zeros((n + 2, n))
    QL[1:-1] = np.eye(n)
    Qb = np.zeros(n + 2)
    if isinstance(bc, _PeriodicPad):
        QL[0, -1] = 1.0
        QL[-1, 0] = 1.0
        return QL, Qb
    QL[0, : len(bc.a_l)] = bc.a_l
    QL[-1, n - len(bc.a_r):] = bc.a_r
    Qb[0], Qb[-1] = bc.b_l, bc.b_r
    return QL, Qb


def _bcs_1d(rng):
    dx = 0.07
    return [
        ("robin1", D.RobinBC((1.0, 0.5, 0.3), (0.7, 0.2, -0.1), dx, 1)),
        ("robin3", D.RobinBC((1.0, 6.0, 10.0), (1.0, 6.0, 10.0), dx, 3)),
        ("dirichlet", D.DirichletBC(-1.0, 2.0)),
        ("dirichlet0", D.Dirichlet0BC(float)),
        ("neumann", D.NeumannBC((0.3, -0.2), dx, 2)),
        ("general", D.GeneralBC([-0.5, 1.0, 0.3, 0.2], [0.4, 1.0, -0.1, 0.6], dx, 2)),
        ("periodic", _PeriodicPad(_abi.BC_PERIODIC)),
    ]


def _ops_1d(rng, n):
    dx = 0.07
    dxv = 0.05 * (1 + 0.5 * rng.random(n + 1))
    return [
        ("c2a2", [D.CenteredDifference(2, 2, dx, n)]),
        ("c4a4", [D.CenteredDifference(4, 4, dx, n, rng.standard_normal(n))]),
        ("c1a6_irregular", [D.CenteredDifference(1, 6, dxv, n)]),
        ("u1a3_mixed", [D.UpwindDifference(1, 3, dx, n, rng.standard_normal(n))]),
        ("u2a2_irregular_neg", [D.UpwindDifference(2, 2, dxv, n, -1.3)]),
        ("sum", [D.CenteredDifference(2, 4, dx, n), D.UpwindDifference(1, 2, dx, n, -0.7)]),
    ]


@pytest.mark.parametrize("sem", [_abi.SEM_BPV, _abi.SEM_LINES])
def test_ghost_is_L_times_QL(hc, rng, sem):
    # Array(A::GhostDerivativeOperator, N) = (Array(L, N) * Q_L, Array(L, N) * Q_b)
    n = 19
    for oname, ops in _ops_1d(rng, n):
        if sem == _abi.SEM_BPV and len(ops) > 1:
            continue
        Lm = sum(padded_matrix(hc, [op], (n + 2,)) for op in ops) if sem == _abi.SEM_LINES \
            else None
        for bname, bc in _bcs_1d(rng):
            A, b = ghost_matrix(hc, ops, {0: bc}, (n,), sem)
            if Lm is not None:
                QL, Qb = q_matrices(bc, n)
                assert rel(A, Lm @ QL) < 1e-13, (oname, bname)
                assert np.abs(b - Lm @ Qb).max() <= 1e-13 * max(np.abs(Lm @ Qb).max(), 1.0), (oname, bname)
            # and column by column against the oracle's apply (unit vectors)
            obc = po.periodic() if isinstance(bc, _PeriodicPad) else po.BC(bc.a_l, bc.b_l, bc.a_r, bc.b_r)
            oops = [_oracle_op(op, 0) for op in ops]
            z = po.ghost_apply(oops, [obc], np.zeros(n), bpv_1d=(sem == _abi.SEM_BPV))
            assert np.abs(b - z).max() <= 1e-13 * max(np.abs(z).max(), 1.0), (oname, bname)
            for j in range(n):
                e = np.zeros(n)
                e[j] = 1.0
                colj = po.ghost_apply(oops, [obc], e, bpv_1d=(sem == _abi.SEM_BPV)) - z
                assert np.abs(A[:, j] - colj).max() <= 1e-12 * max(np.abs(A).max(), 1.0), (oname, bname, j)


def _oracle_op(L, axis):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=axis, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


@pytest.mark.parametrize("msz,N", [((22,), 1), ((22, 3), 1), ((2, 22, 2), 2), ((3, 2, 22), 3),
                                   ((4, 22, 3), 2)])
def test_nd_kron_form(hc, rng, msz, N):
    # Array(L, Mshape) = Kron(I, Kron(Array(L), I)) (concretization.jl:425-444); the reference
    # test concretization.jl (test) :9-131 uses these shapes
    for L in (D.CenteredDifference[N](1, 2, 1.0, 20), D.CenteredDifference[N](3, 3, 1.0, 20),
              D.UpwindDifference[N](1, 2, 1.0, 20, rng.standard_normal(20))):
        A = padded_matrix(hc, [L], msz)
        L1 = padded_matrix(hc, [D.UpwindDifference(1, 2, 1.0, 20, L.coefficients.copy())
                                if L.wind else D.CenteredDifference(L.derivative_order,
                                                                    L.approximation_order, 1.0, 20)],
                           (22,))
        before = int(np.prod(msz[: N - 1]))
        after = int(np.prod(msz[N:]))
        K = np.kron(np.eye(after), np.kron(L1, np.eye(before)))
        assert rel(A, K) < 1e-15


@pytest.mark.parametrize("shape", [(7, 9), (5, 6, 4)])
def test_nd_ghost_columns_match_oracle(hc, rng, shape):
    jsize = tuple(reversed(shape))
    nd = len(shape)
    ops, oops = [], []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](2, 2, dx, n), D.UpwindDifference[N](1, 2, dx, n, -rng.standard_normal(n))]
        oops += [_oracle_op(ops[-2], N - 1), _oracle_op(ops[-1], N - 1)]
    bcs = {0: D.RobinBC((1.0, 0.5, 0.2), (1.0, 0.5, -0.3), 0.05, 2),
           1: D.NeumannBC((0.1, 0.2), 0.05, 1)}
    if nd == 3:
        bcs[2] = _PeriodicPad(_abi.BC_PERIODIC)
    obcs = [po.periodic() if isinstance(bcs[k], _PeriodicPad) else
            po.BC(bcs[k].a_l, bcs[k].b_l, bcs[k].a_r, bcs[k].b_r) for k in range(nd)]
    A, b = ghost_matrix(hc, ops, bcs, jsize, _abi.SEM_LINES)
    z = po.ghost_apply(oops, obcs, np.zeros(shape))
    assert np.abs(b - z.ravel()).max() <= 1e-13 * max(np.abs(z).max(), 1.0)
    n = int(np.prod(shape))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        colj = (po.ghost_apply(oops, obcs, e.reshape(shape)) - z).ravel()
        assert np.abs(A[:, j] - colj).max() <= 1e-12 * np.abs(A).max(), j


def test_periodic_nd_columns(hc, rng):
    # N-D PeriodicBC padding: ghost planes = first / last planes
    shape = (6, 8)
    Lx = D.CenteredDifference[1](2, 2, 0.1, 8)
    Ly = D.CenteredDifference[2](1, 2, 0.1, 6)
    pad = _PeriodicPad(_abi.BC_PERIODIC_ND)
    A, b = ghost_matrix(hc, [Lx, Ly], {0: pad, 1: pad}, (8, 6), _abi.SEM_LINES)
    assert np.all(b == 0)
    oops = [_oracle_op(Lx, 0), _oracle_op(Ly, 1)]
    for j in range(48):
        e = np.zeros(48)
        e[j] = 1.0
        colj = po.ghost_apply(oops, [po.periodic_nd(), po.periodic_nd()], e.reshape(shape)).ravel()
        assert np.abs(A[:, j] - colj).max() <= 1e-12 * np.abs(A).max()
