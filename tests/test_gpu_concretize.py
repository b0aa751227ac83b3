"""GPU tests of the device concretization (dop_plan_csr_rows / dop_plan_csr_fill) and the
banded direct solve (dop_solver_*), through the C ABI.

Parity: the exported matrix equals, column by column, the FP64 oracle's apply on unit vectors
(A[:, j] = oracle(e_j) - oracle(0), b = oracle(0)) to 1e-12 of max|A|, equals the CPU build of
the same enumeration (tests/hostcheck) bit for bit, and the reference tests' golden matrices
(`Array(L) == correct`).  At larger sizes: A_l * u + A_b equals the fused apply (FP64, 1e-12).
Solves: poisson.jl (the reference's own test: exact for a quadratic), agreement with a dense
LU of the oracle's matrix to max(1e-12, 1e-14 cond_inf(A)) (both are backward stable), backward
error |A x + b - u| / (|A| |x|) below 1e-13, and
DOP_ESINGULAR on an exactly singular operator."""
import ctypes as C
import os
import time

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import Plan, _PeriodicPad, make_desc
from diffeqoperators_amd.concretize import Solver, plan_csr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC_PATH = os.environ.get("DOP_HOSTCHECK_LIBRARY") or os.path.join(ROOT, "tests", "hostcheck", "_build",
                                                                "libdop_hostcheck.so")


def oop(L):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=L.N - 1, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


def obc(bc):
    if isinstance(bc, _PeriodicPad):
        return po.periodic() if bc.kind == _abi.BC_PERIODIC else po.periodic_nd()
    return po.BC(bc.a_l, bc.b_l, bc.a_r, bc.b_r)


def dense(rp, col, val, n_rows, n_cols):
    rp, col, val = rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    A = np.zeros((n_rows, n_cols))
    for i in range(n_rows):
        A[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return A


def ghost_case(name, rng):
    """(ops, bcs {axis: bc}, julia size, semantics)"""
    dx = 0.07
    if name == "1d_robin_c2":
        return [D.CenteredDifference(2, 2, dx, 31)], {0: D.RobinBC((1.0, 0.5, 0.3), (0.7, 0.2, -0.1), dx, 2)}, (31,), _abi.SEM_BPV
    if name == "1d_general_c4":
        return ([D.CenteredDifference(4, 4, dx, 29, rng.standard_normal(29))],
                {0: D.GeneralBC([-0.5, 1.0, 0.3, 0.2], [0.4, 1.0, -0.1, 0.6], dx, 2)}, (29,), _abi.SEM_BPV)
    if name == "1d_periodic_upwind":
        return ([D.UpwindDifference(1, 3, dx, 33, rng.standard_normal(33))],
                {0: _PeriodicPad(_abi.BC_PERIODIC)}, (33,), _abi.SEM_BPV)
    if name == "1d_sum_neumann":
        return ([D.CenteredDifference(2, 4, dx, 40), D.UpwindDifference(1, 2, dx, 40, -0.7)],
                {0: D.NeumannBC((0.3, -0.2), dx, 1)}, (40,), _abi.SEM_LINES)
    if name == "1d_general_c2_upwind":
        return ([D.CenteredDifference(2, 4, dx, 37), D.UpwindDifference(1, 3, dx, 37, rng.standard_normal(37))],
                {0: D.GeneralBC([-0.5, 1.0, 0.3], [0.4, 1.0, -0.1], dx, 2)}, (37,), _abi.SEM_BPV)
    if name == "2d_robin_sum":
        n1, n2 = 13, 11
        dxv = 0.05 * (1 + 0.5 * rng.random(n2 + 1))
        ops = [D.CenteredDifference[1](2, 4, dx, n1), D.CenteredDifference[2](2, 2, dxv, n2),
               D.UpwindDifference[1](1, 3, dx, n1, -rng.standard_normal(n1))]
        return ops, {0: D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), dx), 1: D.Dirichlet0BC(float)}, (n1, n2), _abi.SEM_LINES
    if name == "3d_mixed":
        s = (7, 6, 5)
        ops = [D.CenteredDifference[1](2, 2, dx, 7), D.CenteredDifference[2](2, 2, dx, 6),
               D.UpwindDifference[3](1, 2, dx, 5, rng.standard_normal(5))]
        return ops, {0: D.DirichletBC(1.0, -2.0), 1: D.NeumannBC((0.1, 0.2), dx, 2),
                     2: _PeriodicPad(_abi.BC_PERIODIC)}, s, _abi.SEM_LINES
    raise KeyError(name)


GHOST = ["1d_robin_c2", "1d_general_c4", "1d_periodic_upwind", "1d_sum_neumann", "1d_general_c2_upwind",
         "2d_robin_sum", "3d_mixed"]


def export(plan):
    rp, col, val, b, kl, ku = plan_csr(plan, torch.device("cuda"))
    torch.cuda.synchronize()
    return rp, col, val, b, kl, ku


@pytest.mark.parametrize("name", GHOST)
def test_ghost_export_matches_oracle_columns(rng, name):
    ops, bcs, jsize, sem = ghost_case(name, rng)
    print(f"[concretize] {name} {jsize}", flush=True)
    d, keep = make_desc(ops, jsize, _abi.F64, input_kind=_abi.INPUT_GHOST, semantics=sem, bcs=bcs)
    plan = Plan(d, keep + [bcs])
    rp, col, val, b, kl, ku = export(plan)
    n = int(np.prod(jsize))
    A = dense(rp, col, val, n, n)
    shape = tuple(reversed(jsize))
    oops = [oop(L) for L in ops]
    obcs = [obc(bcs[k]) for k in range(len(jsize))]
    bpv = sem == _abi.SEM_BPV
    z = po.ghost_apply(oops, obcs, np.zeros(shape), bpv_1d=bpv).ravel()
    assert np.abs(b.cpu().numpy() - z).max() <= 1e-12 * max(np.abs(z).max(), 1.0)
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cj = po.ghost_apply(oops, obcs, e.reshape(shape), bpv_1d=bpv).ravel() - z
        assert np.abs(A[:, j] - cj).max() <= 1e-12 * np.abs(A).max(), j
    # bandwidths reported by the export
    r, c = np.nonzero(A)
    assert kl == max(0, int((r - c).max())) and ku == max(0, int((c - r).max()))
    # bit for bit the CPU build of the same enumeration, when it is present
    if os.path.exists(HC_PATH):
        hc = C.CDLL(HC_PATH)
        hc.hc_csr.argtypes = [C.POINTER(_abi.DopDesc)] + [C.c_void_p] * 4 + [C.POINTER(C.c_int64)]
        hrp = np.zeros(n + 1, dtype=np.int64)
        nnz = C.c_int64()
        assert hc.hc_csr(C.byref(d), hrp.ctypes.data, None, None, None, C.byref(nnz)) == 0
        hcol = np.zeros(nnz.value, dtype=np.int64)
        hval = np.zeros(nnz.value)
        hb = np.zeros(n)
        assert hc.hc_csr(C.byref(d), hrp.ctypes.data, hcol.ctypes.data, hval.ctypes.data,
                         hb.ctypes.data, C.byref(nnz)) == 0
        assert np.array_equal(hrp, rp.cpu().numpy())
        assert np.array_equal(hcol, col.cpu().numpy())
        assert np.array_equal(hval.view(np.int64), val.cpu().numpy().view(np.int64))
        assert np.array_equal(hb.view(np.int64), b.cpu().numpy().view(np.int64))


def test_array_of_operator_matches_golden(fixtures):
    for e in fixtures["centered_uniform"]:
        L = D.CenteredDifference(e["d"], e["a"], e["dx"], e["N"])
        A = D.Array(L).cpu().numpy()
        assert np.abs(A - np.array(e["A"])).max() <= 1e-14 * np.abs(A).max(), e["name"]
    for e in fixtures["upwind_uniform"]:
        n = len(e["A"])
        coeff = e["coeff"] if np.ndim(e["coeff"]) == 0 else np.array(e["coeff"])
        L = D.UpwindDifference(e["d"], e["a"], e["dx"], n, coeff)
        S = D.sparse(L)
        assert S.layout == torch.sparse_csr and tuple(S.shape) == (n, n + 2)
        A = S.to_dense().cpu().numpy()
        assert np.abs(A - np.array(e["A"])).max() <= 1e-13 * np.abs(A).max(), e["name"]


def test_nd_operator_array_times_vec_equals_apply(rng):
    # concretization.jl (test) :9-131: vec(L * M) == Array(L, size(M)) * vec(M)
    for N, msz in ((1, (22, 2, 2)), (2, (2, 22, 2)), (3, (3, 2, 22))):
        L = D.CenteredDifference[N](3, 3, 1.0, 20)
        M = torch.from_numpy(rng.standard_normal(tuple(reversed(msz)))).cuda()
        A = D.Array(L, msz).cpu().numpy()
        lm = (L * M).cpu().numpy()
        got = A @ M.cpu().numpy().ravel()
        assert np.abs(got - lm.ravel()).max() <= 1e-12 * np.abs(lm).max()


@pytest.mark.parametrize("shape", [(256, 256), (40, 48, 56)])
def test_export_times_u_equals_apply_large(rng, shape):
    import scipy.sparse as sp
    nd = len(shape)
    jsize = tuple(reversed(shape))
    ops = []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](2, 4, dx, n), D.UpwindDifference[N](1, 3, dx, n, rng.standard_normal(n))]
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.2), (1.0, 0.5, -0.1), [0.05] * nd, 1, jsize))
    A = sum(ops[1:], ops[0]) * Q
    print(f"[concretize] large {shape}", flush=True)
    Al, Ab = D.sparse(A, jsize)
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).cuda()
    du = (A * u).cpu().numpy().ravel()
    S = sp.csr_matrix((Al.values().cpu().numpy(), Al.col_indices().cpu().numpy(),
                       Al.crow_indices().cpu().numpy()), shape=tuple(Al.shape))
    got = S @ u.cpu().numpy().ravel() + Ab.cpu().numpy()
    assert np.abs(got - du).max() <= 1e-12 * np.abs(du).max()


# ---------------------------------------------------------------------------------------
# A \ u

def test_poisson_reference_example():
    # poisson.jl: Δu = f on (0,1), u(0) = a, u(1) = b; 2nd order exact for the quadratic
    f, a, b = 1.0, -1.0, 2.0
    nknots = 10
    h = 1.0 / (nknots + 1)
    Δ = D.CenteredDifference(2, 2, h, nknots)
    bc = D.DirichletBC(a, b)
    u = D.ldiv(Δ * bc, torch.full((nknots,), f, dtype=torch.float64, device="cuda"))
    knots = h * np.arange(1, nknots + 1)
    ana = f / 2 * knots ** 2 + (b - a - f / 2) * knots + a
    assert np.abs(u.cpu().numpy() - ana).max() <= 1e-12 * np.abs(ana).max()


def _oracle_dense(ops, bcs, jsize, sem):
    shape = tuple(reversed(jsize))
    oops = [oop(L) for L in ops]
    obcs = [obc(bcs[k]) for k in range(len(jsize))]
    bpv = sem == _abi.SEM_BPV
    z = po.ghost_apply(oops, obcs, np.zeros(shape), bpv_1d=bpv).ravel()
    n = int(np.prod(jsize))
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = po.ghost_apply(oops, obcs, e.reshape(shape), bpv_1d=bpv).ravel() - z
    return A, z


@pytest.mark.parametrize("name", ["1d_robin_c2", "1d_general_c2_upwind", "2d_robin_sum", "3d_mixed"])
def test_solve_matches_dense_lu(rng, name):
    ops, bcs, jsize, sem = ghost_case(name, rng)
    A = sum(ops[1:], ops[0]) if len(ops) > 1 else ops[0]
    if len(jsize) == 1:
        Q = bcs[0] if not isinstance(bcs[0], _PeriodicPad) else D.PeriodicBC(float)
    else:
        Q = D.compose(*[D.MultiDimBC[k + 1](bcs[k], jsize) if not isinstance(bcs[k], _PeriodicPad)
                        else D.PeriodicBC[k + 1](float, jsize) for k in range(len(jsize))])
    G = A * Q
    shape = tuple(reversed(jsize))
    print(f"[solve] {name} {jsize}", flush=True)
    u = rng.uniform(-1, 1, shape)
    x = D.ldiv(G, torch.from_numpy(u).cuda()).cpu().numpy().ravel()
    # the public path takes the BoundaryPaddedVector kernels for every 1-D Q*u
    Ad, z = _oracle_dense(ops, bcs, jsize, _abi.SEM_BPV if len(jsize) == 1 else sem)
    ref = np.linalg.solve(Ad, u.ravel() - z)
    cond = np.linalg.cond(Ad, np.inf)
    print(f"[solve] cond_inf = {cond:.3g}", flush=True)
    assert cond < 1e8     # well posed, so that the forward comparison means something
    assert np.abs(x - ref).max() <= max(1e-12, 1e-14 * cond) * np.abs(ref).max()
    berr = np.abs(Ad @ x + z - u.ravel()).max() / (np.abs(Ad).sum(axis=1).max() * np.abs(x).max())
    assert berr < 1e-13


def test_solve_batched_rhs_equals_single(rng):
    n = 200
    G = (D.CenteredDifference(2, 4, 0.01, n) + D.UpwindDifference(1, 2, 0.01, n, rng.standard_normal(n))) \
        * D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.2), 0.01, 2)
    S = Solver(G, (n,))
    R = torch.from_numpy(rng.standard_normal((5, n))).cuda()
    X = S.solve(R)
    for i in range(5):
        xi = S.solve(R[i].contiguous())
        assert torch.equal(X[i], xi)


def test_solve_singular_raises():
    # centered first derivative with homogeneous Dirichlet on an odd number of points: the
    # skew-symmetric tridiagonal matrix is singular and its LU meets an exact zero pivot
    G = D.CenteredDifference(1, 2, 1.0, 5) * D.Dirichlet0BC(float)
    with pytest.raises(_abi.DopError) as e:
        D.ldiv(G, torch.ones(5, dtype=torch.float64, device="cuda"))
    assert e.value.code == _abi.ESINGULAR


def test_solve_errors():
    L = D.CenteredDifference(2, 2, 0.1, 10)
    with pytest.raises(TypeError):
        D.ldiv(L, torch.ones(12, dtype=torch.float64, device="cuda"))    # not L*Q
    G = L * D.Dirichlet0BC(float)
    with pytest.raises(TypeError):
        D.ldiv(G, torch.ones(10, dtype=torch.float32, device="cuda"))
    with pytest.raises(TypeError):
        D.ldiv(G, torch.ones(10, dtype=torch.float64))                    # no CPU path


def test_solve_large_1d_backward_error(rng):
    # a 65536-point heat operator: the solve inverts the fused apply (size-independent check)
    n = 65536
    h = 1.0 / (n + 1)
    G = D.CenteredDifference(2, 2, h, n) * D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), h, 1)
    u = torch.from_numpy(rng.uniform(-1, 1, n)).cuda()
    print("[solve] 1-D 65536", flush=True)
    t0 = time.perf_counter()
    S = Solver(G, (n,))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x = S.solve(u)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"[solve] factor {1e3 * (t1 - t0):.1f} ms (incl. export), solve {1e3 * (t2 - t1):.1f} ms, "
          f"kl={S.kl} ku={S.ku}", flush=True)
    r = (G * x) - u
    Anorm = 4.0 / h ** 2
    berr = r.abs().max().item() / (Anorm * x.abs().max().item())
    assert berr < 1e-13
