"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the same seeded
inputs.  Tolerances (normwise, max|got - ref| / max|ref|): 1e-12 for FP64 storage; for
FP32 / BF16 storage the oracle runs on the same quantized input and the bounds are 1e-5
(FP32) and 1e-2 (BF16, whose output rounding alone is 2^-9)."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.bfloat16: 1e-2}


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    po.set_threads(min(16, os.cpu_count() or 1))
    yield
    po.set_threads(1)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oop(L):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=L.N - 1, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


def obc(bc):
    return po.BC(bc.a_l, bc.b_l, bc.a_r, bc.b_r)


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda").to(dtype)


def host64(t):
    return t.to(torch.float64).cpu().numpy()


# ---------------------------------------------------------------------------------------
# workload 1: 1-D heat-equation RHS (heat_equation.jl:17-30)

def test_heat_equation_rhs_n100():
    n = 100
    h = 1.0 / (n + 1)
    x = h * np.arange(1, n + 1)
    rng = np.random.default_rng(0)
    u = np.sin(2 * np.pi * x) + rng.normal(0, 1e-3, n)
    Δ = D.CenteredDifference(2, 2, h, n)
    bc = D.Dirichlet0BC(float)
    print("case: heat N=100 f64", flush=True)
    got = host64(Δ * bc * to_dev(u))
    ref = po.ghost_apply([oop(Δ)], [obc(bc)], u)
    assert rel(got, ref) < 1e-12
    # mul! form into an existing buffer
    du = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    D.mul_(du, Δ * bc, to_dev(u))
    assert rel(host64(du), ref) < 1e-12


@pytest.mark.parametrize("d,a", [(1, 2), (2, 2), (2, 4), (3, 2), (4, 4), (2, 6), (4, 8), (3, 8)])
@pytest.mark.parametrize("coeff", ["const", "vector"])
def test_1d_centered(d, a, coeff):
    rng = np.random.default_rng(d * 10 + a)
    n = 257
    u = rng.standard_normal(n)
    c = 1.0 if coeff == "const" else rng.standard_normal(n)
    L = D.CenteredDifference(d, a, 0.01, n, c)
    bc = D.RobinBC((1.0, 0.5, 0.2), (0.8, 0.3, -0.4), 0.01, 2)
    print(f"case: 1d centered d={d} a={a} {coeff}", flush=True)
    got = host64((L * bc) * to_dev(u))
    ref = po.ghost_apply([oop(L)], [obc(bc)], u, bpv_1d=True)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("d,a,off", [(1, 1, 0), (1, 2, 0), (1, 3, 0), (2, 2, 0), (2, 3, 1),
                                     (1, 4, 2), (3, 2, 0)])
@pytest.mark.parametrize("grid", ["uniform", "irregular"])
@pytest.mark.parametrize("wind", ["pos", "neg", "mixed", "scaled"])
def test_1d_upwind(d, a, off, grid, wind):
    rng = np.random.default_rng(d * 100 + a * 10 + off)
    n = 301
    dx = 0.01 if grid == "uniform" else 0.005 * (1 + 0.5 * rng.random(n + 1))
    c = {"pos": 1.3, "neg": -0.7, "mixed": rng.standard_normal(n), "scaled": 1.0}[wind]
    L = D.UpwindDifference(d, a, dx, n, c, offside=off)
    if wind == "scaled":
        L = -2.0 * L
    bc = D.NeumannBC((0.3, -0.2), dx, 1)
    u = rng.standard_normal(n)
    print(f"case: 1d upwind d={d} a={a} off={off} {grid} {wind}", flush=True)
    got = host64((L * bc) * to_dev(u))
    ref = po.ghost_apply([oop(L)], [obc(bc)], u, bpv_1d=True)
    assert rel(got, ref) < 1e-12


# ---------------------------------------------------------------------------------------
# workload 2: 2-D uniform Laplacian, Robin(1, 0.5, 0) on every edge, FP64

def laplacian_robin(n, h):
    Lx = D.CenteredDifference[1](2, 4, h, n)
    Ly = D.CenteredDifference[2](2, 4, h, n)
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), (h, h), 1, (n, n)))
    return (Lx + Ly) * Q, [Lx, Ly], Q


@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_2d_laplacian_robin_f64(n):
    h = 1.0 / (n + 1)
    A, ops, Q = laplacian_robin(n, h)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, (n, n))
    print(f"case: 2d laplacian robin n={n} f64", flush=True)
    got = host64(A * to_dev(u))
    ref = po.ghost_apply([oop(L) for L in ops], [obc(b.bc) for b in Q.bcs], u)
    assert rel(got, ref) < 1e-12


def test_2d_laplacian_linearity_full_size():
    # a size-independent property at the full 4096^2 size: the operator (gamma = 0) is linear
    n = 4096
    A, _, _ = laplacian_robin(n, 1.0 / (n + 1))
    g = torch.Generator(device="cuda").manual_seed(7)
    u = torch.rand((n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    v = torch.rand((n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    lhs = A * (0.75 * u - 1.5 * v)
    rhs = 0.75 * (A * u) - 1.5 * (A * v)
    err = ((lhs - rhs).abs().max() / rhs.abs().max()).item()
    assert err < 1e-13


# ---------------------------------------------------------------------------------------
# workload 3: 3-D non-uniform CenteredDifference(2,6) + UpwindDifference(1,3) per axis

def nonuniform_3d(n, seed=10):
    h = 1.0 / (n + 1)
    ops = []
    for k in (1, 2, 3):
        r = np.random.default_rng(seed + k)
        dx = h * (1 + 0.5 * r.random(n + 1))
        v = np.random.default_rng(20 + k).standard_normal(n)
        ops += [D.CenteredDifference[k](2, 6, dx, n), D.UpwindDifference[k](1, 3, dx, n, -v)]
    Q = D.compose(*D.Dirichlet0BC(float, (n, n, n)))
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    return A * Q, ops, Q


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("n", [48, 160])
def test_3d_nonuniform(dtype, n):
    A, ops, Q = nonuniform_3d(n)
    rng = np.random.default_rng(3)
    u64 = rng.uniform(-1, 1, (n, n, n))
    ud = to_dev(u64, dtype)
    uq = host64(ud)                       # quantized input, exactly representable in f64
    print(f"case: 3d nonuniform n={n} {dtype}", flush=True)
    got = host64(A * ud)
    ref = po.ghost_apply([oop(L) for L in ops], [obc(b.bc) for b in Q.bcs], uq)
    err = rel(got, ref)
    print(f"  max rel err vs FP64 oracle: {err:.3e}", flush=True)
    assert err < TOL[dtype]


# ---------------------------------------------------------------------------------------
# padded arrays: L*M, mul!(du, L, M; overwrite), composite mul!

def test_padded_single_operator_and_accumulate():
    rng = np.random.default_rng(5)
    n1, n2 = 300, 77
    L = D.CenteredDifference[1](4, 4, 0.1, n1, rng.standard_normal(n1))
    M = rng.standard_normal((n2, n1 + 2))
    got = host64(L * to_dev(M))
    ref = po.mul_nd(oop(L), M, (n2, n1))
    assert rel(got, ref) < 1e-12
    # trimmed transverse padding + overwrite=false
    M2 = rng.standard_normal((n2 + 2, n1 + 2))
    du0 = rng.standard_normal((n2, n1))
    du = to_dev(du0)
    D.mul_(du, L, to_dev(M2), overwrite=False)
    ref2 = po.mul_nd(oop(L), M2, (n2, n1)) + du0
    assert rel(host64(du), ref2) < 1e-12


@pytest.mark.parametrize("nd", [2, 3])
def test_composite_mul_padded(nd):
    rng = np.random.default_rng(6)
    n = 40 if nd == 3 else 200
    ops = [D.CenteredDifference[k](2, 4, 0.5, n) for k in range(1, nd + 1)]
    ops.append(D.UpwindDifference[2](1, 2, 0.5, n, rng.standard_normal(n)))
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    M = rng.standard_normal((n + 2,) * nd)
    du = torch.empty((n,) * nd, dtype=torch.float64, device="cuda")
    D.mul_(du, A, to_dev(M))
    ref = po.composite_mul_padded([oop(L) for L in ops], M)
    assert rel(host64(du), ref) < 1e-12


def test_ghost_accumulate_and_bad_inputs():
    n = 128
    A, ops, Q = laplacian_robin(n, 1.0 / (n + 1))
    rng = np.random.default_rng(8)
    u = rng.uniform(-1, 1, (n, n))
    du0 = rng.standard_normal((n, n))
    du = to_dev(du0)
    D.mul_(du, A, to_dev(u), overwrite=False)
    ref = po.ghost_apply([oop(L) for L in ops], [obc(b.bc) for b in Q.bcs], u) + du0
    assert rel(host64(du), ref) < 1e-12
    ud = to_dev(u)
    with pytest.raises(ValueError):
        D.mul_(ud, A, ud)                         # aliasing
    with pytest.raises(AssertionError):
        A * to_dev(rng.standard_normal((n, n + 1)))   # BC sized for (n, n)


# ---------------------------------------------------------------------------------------
# PeriodicBC: line wrap in 1-D (bc_operators.jl:192), first/last planes for an all-periodic
# N-D BC (multi_dim_bc_operators.jl:221-252), wrap per line when composed with other BCs

def obc_any(Q, ndim, jsize):
    from diffeqoperators_amd.apply import _PeriodicPad, _bcs_of
    bcs, _ = _bcs_of(Q, ndim, jsize)
    out = []
    for k in range(ndim):
        b = bcs[k]
        if isinstance(b, _PeriodicPad):
            out.append(po.periodic() if b.kind == _abi.BC_PERIODIC else po.periodic_nd())
        else:
            out.append(obc(b))
    return out


def test_periodic_1d_second_order_check():
    h = 0.025 * np.pi
    N = int(round(2 * (np.pi / h)))
    x = -np.pi + h * np.arange(N)
    u0 = np.cos(x)
    L2 = D.CenteredDifference(2, 2, h, N)
    Q = D.PeriodicBC(float)
    print("case: periodic 1-D second-order check", flush=True)
    got = host64(L2 * Q * to_dev(u0))
    M = np.diag(np.full(N, -2.0)) + np.diag(np.ones(N - 1), 1) + np.diag(np.ones(N - 1), -1)
    M[-1, 0] = M[0, -1] = 1.0
    assert rel(got, (M / (h * h)) @ u0) < 1e-12
    assert rel(got, po.ghost_apply([oop(L2)], [po.periodic()], u0, bpv_1d=True)) < 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("form", ["all", "mixed"])
@pytest.mark.parametrize("shape", [(1000, 1000), (96, 80, 72)])
def test_periodic_nd(dtype, form, shape):
    rng = np.random.default_rng(77)
    js = tuple(reversed(shape))
    nd = len(shape)
    if form == "all":
        Q = D.compose(*D.PeriodicBC(float, js))
    else:
        Q = D.compose(D.PeriodicBC[1](float, js), *[D.Dirichlet0BC[k](float, js) for k in range(2, nd + 1)])
    ops = [D.CenteredDifference[N](2, 4, 0.01, js[N - 1]) for N in range(1, nd + 1)]
    ops.append(D.UpwindDifference[1](1, 2, 0.01, js[0], -rng.standard_normal(js[0])))
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    A = A * Q
    u = to_dev(rng.uniform(-1, 1, shape), dtype)
    print(f"case: periodic {form} {shape} {dtype}", flush=True)
    got = host64(A * u)
    ref = po.ghost_apply([oop(L) for L in ops], obc_any(Q, nd, js), host64(u))
    assert rel(got, ref) < TOL[dtype]


# ---------------------------------------------------------------------------------------
# GeneralBC (bc_operators.jl:85-170), as in KdV.jl:28: third derivative with a BC on u, u', u'''

@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_general_bc_kdv(dtype):
    n = 400
    dx = 20.0 / (n + 1)
    rng = np.random.default_rng(31)
    bc = D.GeneralBC([0.0, 1.0, -0.6, 0.0, -1.0], [0.0, 1.0, -0.4, 0.0, -1.0], dx, 3)
    L3 = D.CenteredDifference(3, 2, dx, n)
    L1 = D.UpwindDifference(1, 1, dx, n, -rng.standard_normal(n))
    A = (L3 + L1) * bc
    u = rng.uniform(-1, 1, n)
    print(f"case: GeneralBC KdV 1-D {dtype}", flush=True)
    ud = to_dev(u, dtype)
    got = host64(A * ud)
    ref = po.ghost_apply([oop(L3), oop(L1)], [po.general([0.0, 1.0, -0.6, 0.0, -1.0],
                                                        [0.0, 1.0, -0.4, 0.0, -1.0], dx, 3)],
                         host64(ud), bpv_1d=True)
    assert rel(got, ref) < TOL[dtype]


def test_general_bc_2d():
    n = 300
    h = 1.0 / (n + 1)
    rng = np.random.default_rng(32)
    atom = D.GeneralBC([0.3, 1.0, 0.5, 0.1], [-0.2, 1.0, -0.5, 0.2], h, 2)
    Q = D.compose(*D.MultiDimBC(atom, (n, n)))
    ops = [D.CenteredDifference[1](2, 4, h, n), D.CenteredDifference[2](2, 4, h, n)]
    A = (ops[0] + ops[1]) * Q
    u = rng.uniform(-1, 1, (n, n))
    print("case: GeneralBC 2-D f64", flush=True)
    got = host64(A * to_dev(u))
    ref = po.ghost_apply([oop(L) for L in ops], [obc(atom), obc(atom)], u)
    assert rel(got, ref) < 1e-12


# ---------------------------------------------------------------------------------------
# Q*u as a lazy padded object (boundary_padded_arrays.jl:14-180): L*(Q*u) is the fused ghost
# apply; materialize() is the padded array, whose plain padded apply agrees

@pytest.mark.parametrize("shape", [(200,), (90, 70), (20, 24, 28)])
def test_boundary_padded_lazy_and_materialized(shape):
    rng = np.random.default_rng(41)
    js = tuple(reversed(shape))
    nd = len(shape)
    atom = D.RobinBC((1.0, 0.5, 0.3), (1.0, 0.5, -0.2), 0.01, 2)
    if nd == 1:
        Q = atom
        L = D.CenteredDifference(2, 4, 0.01, js[0])
        ops = [L]
    else:
        Q = D.compose(*D.MultiDimBC(atom, js))
        ops = [D.CenteredDifference[N](2, 4, 0.01, js[N - 1]) for N in range(1, nd + 1)]
    A = ops[0]
    for op in ops[1:]:
        A = A + op
    u = to_dev(rng.uniform(-1, 1, shape))
    print(f"case: Q*u lazy/materialized {shape}", flush=True)
    lazy = A * (Q * u)
    fused = (A * Q) * u
    assert torch.equal(lazy, fused)
    du = torch.empty_like(u)
    D.mul_(du, A * Q, u)
    assert torch.equal(du, fused)
    # the materialized padded array: ghosts as the oracle pads each line
    M = (Q * u).materialize()
    Mh = host64(M)
    uh = host64(u)
    ob = obc(atom)
    line = np.moveaxis(uh, -1, 0).reshape(js[0], -1).T[0]
    gl, gr = ob.pad(line)
    first = np.moveaxis(Mh, -1, 0)
    sl = (0,) + tuple(1 for _ in range(nd - 1))
    sr = (js[0] + 1,) + tuple(1 for _ in range(nd - 1))
    assert abs(first[sl] - gl) < 1e-14 and abs(first[sr] - gr) < 1e-14
    if nd == 1:
        # 1-D padded apply (line kernels) against the ghost apply (BPV kernels): the same
        # numbers for constant coefficients
        assert rel(host64(L * M), host64(fused)) < 1e-13
    else:
        # corners of a composed padding are zero (boundary_padded_arrays.jl:158-180)
        corner = tuple(0 for _ in range(nd))
        assert Mh[corner] == 0.0
