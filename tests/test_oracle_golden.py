"""Pins the FP64 oracle (oracle/dop_oracle.c) before it is trusted as the checker:
against the analytic matrices and values held by DiffEqOperators.jl's own tests
(tests/golden/reference_fixtures.json), and against what the mathematics fixes
(polynomial exactness, closed-form Robin ghosts, invariants)."""
import math

import numpy as np
import pytest

import pyoracle as po


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


# ---------------------------------------------------------------------------------------
# golden matrices from the reference tests

def test_centered_uniform_matrices(fixtures):
    # derivative_operators_interface.jl:121-143, 284-295; convolutions.jl (test) :45-81
    for e in fixtures["centered_uniform"]:
        op = po.Op(0, e["d"], e["a"], e["dx"], e["N"])
        A = np.array(e["A"])
        assert rel(op.matrix(), A) < 1e-14, e["name"]
        assert rel(op.matrix(bpv=True), A) < 1e-14, e["name"]


def test_centered_irregular_matrices(fixtures):
    # derivative_operators_interface.jl:229-281 (rows 2:end-1 for the 4-point stencils)
    irr = fixtures["centered_irregular"]
    for e in irr["entries"]:
        op = po.Op(0, e["d"], e["a"], np.array(irr["dx"]), len(irr["x"]) - 2)
        M = op.matrix()
        if e["rows"] == "interior":
            M = M[1:-1]
        assert rel(M, np.array(e["A"])) < 1e-13, e["name"]


@pytest.mark.parametrize("group", ["upwind_uniform", "upwind_irregular"])
def test_upwind_matrices(fixtures, group):
    # upwind_operators_interface.jl:152-839
    for e in fixtures[group]:
        n = len(e["A"])
        dx = e["dx"] if np.ndim(e["dx"]) == 0 else np.array(e["dx"])
        op = po.Op(1, e["d"], e["a"], dx, n, coeff=e["coeff"])
        A = np.array(e["A"])
        assert rel(op.matrix(), A) < 1e-13, e["name"]
        assert rel(op.matrix(bpv=True), A) < 1e-13, e["name"]


def test_stencil_weights(fixtures):
    # derivative_operators_interface.jl:146-161 (centered), :203-214 (upwind)
    cw = fixtures["centered_weights"]
    for key, a in (("a2", 2), ("a4", 4)):
        for d in (1, 2, 3):
            op = po.Op(0, d, a, 1.0, 20)
            assert np.allclose(op.stencil, cw[key][d - 1], atol=1e-10, rtol=0)
    uw = fixtures["upwind_weights"]
    for key, a in (("a1", 1), ("a2", 2)):
        for j, d in enumerate((1, 3)):
            op = po.Op(1, d, a, 1.0, 20)
            assert np.allclose(op.stencil, uw[key][j], atol=1e-10, rtol=0)


def test_robin_third_order(fixtures):
    # robin.jl:92-100: ghosts -0.4 and 125/12
    r = fixtures["robin3"]
    bc = po.robin(r["l"], r["r"], r["dx"], r["order"])
    gl, gr = bc.pad(np.array(r["u"], dtype=float))
    assert math.isclose(gl, r["ghost_l"], rel_tol=1e-13)
    assert math.isclose(gr, r["ghost_r"], rel_tol=1e-13)


def test_robin_closed_form(rng):
    # robin.jl:13-43 (uniform) and :46-90 (vector dx with equal spacings)
    for i in range(1, 6):
        al, bl, cl, ar, br, cr, dx = rng.random(7) + 0.05
        n = 5 * i
        u = rng.random(n)
        exp_l = (cl - (bl / dx) * u[0]) / (al - bl / dx)
        exp_r = (cr + (br / dx) * u[-1]) / (ar + br / dx)
        for dxa in (dx, dx * np.ones(n)):
            bc = po.robin((al, bl, cl), (ar, br, cr), dxa)
            gl, gr = bc.pad(u)
            assert math.isclose(gl, exp_l, rel_tol=1e-12)
            assert math.isclose(gr, exp_r, rel_tol=1e-12)


def test_dirichlet_is_robin_with_signed_zero_taps():
    # bc_operators.jl:176-178: a = -s[2:end] / Inf -> +-0.0, b = gamma
    bc = po.dirichlet(0.3, -1.7)
    assert np.all(bc.a_l == 0.0) and np.all(bc.a_r == 0.0)
    assert bc.b_l == 0.3 and bc.b_r == -1.7
    assert len(bc.a_l) == 2


# ---------------------------------------------------------------------------------------
# what the mathematics fixes

@pytest.mark.parametrize("d,a", [(1, 2), (2, 2), (2, 4), (3, 2), (4, 4), (2, 6), (4, 8), (3, 6)])
def test_centered_polynomial_exactness(d, a):
    # interior rows differentiate polynomials of degree <= d+a-1 exactly (uniform grid)
    n, h = 30, 0.1
    op = po.Op(0, d, a, h, n)
    x = h * np.arange(n + 2)
    for deg in range(d + a):
        f = x ** deg
        df = (math.factorial(deg) / math.factorial(deg - d) * x[1:-1] ** (deg - d)) if deg >= d else 0 * x[1:-1]
        out = op.mul_padded(f)
        scale = max(np.abs(f).max() / h ** d, 1.0)
        assert np.abs(out - df).max() / scale < 1e-9, (d, a, deg)


@pytest.mark.parametrize("d,a", [(1, 2), (2, 2), (2, 4), (4, 2)])
def test_nonuniform_polynomial_exactness(rng, d, a):
    # every row (interior and boundary) of the non-uniform operator differentiates
    # polynomials of degree d .. SL-1 exactly (each row has at least SL points)
    n = 25
    dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
    op = po.Op(0, d, a, dx, n)
    x = np.concatenate([[0.0], np.cumsum(dx)])
    for deg in range(d, op.SL):
        f = x ** deg
        df = math.factorial(deg) / math.factorial(deg - d) * x[1:-1] ** (deg - d)
        out = op.mul_padded(f)
        scale = max(np.abs(f).max() / dx.min() ** d, 1.0)
        assert np.abs(out - df).max() / scale < 1e-9, (d, a, deg)


def test_weights_sum_to_zero():
    for d in (1, 2, 3, 4):
        for n in range(d + 1, 12):
            w = po.calculate_weights(d, 0.37, np.linspace(0.0, 1.0, n))
            assert abs(w.sum()) < 1e-12 * np.abs(w).sum()


def test_fornberg_textbook():
    assert np.allclose(po.calculate_weights(2, 0.0, [-1, 0, 1]), [1, -2, 1], atol=0)
    assert np.allclose(po.calculate_weights(1, 0.0, [-2, -1, 0, 1, 2]),
                       [1 / 12, -2 / 3, 0, 2 / 3, -1 / 12], atol=1e-15)
    assert np.allclose(po.calculate_weights(0, 0.5, [0, 1]), [0.5, 0.5], atol=1e-15)


def test_uniform_high_boundary_is_reflected_low():
    # derivative_operator.jl:114-116
    for d, a in [(2, 4), (3, 4), (4, 4), (1, 6)]:
        op = po.Op(0, d, a, 0.3, 20)
        B, L = op.BPC, op.BSL
        low = op.low[: B * L].reshape(B, L)
        high = op.high[: B * L].reshape(B, L)
        assert np.array_equal(high, ((-1) ** d) * low[::-1, ::-1])


@pytest.mark.parametrize("d,a,off", [(1, 2, 1), (1, 4, 2), (1, 6, 3), (2, 3, 2), (3, 4, 3)])
def test_max_offside_upwind_equals_centered(rng, d, a, off):
    # concretization.jl (test) :134-200
    M = rng.random(22)
    L = po.Op(0, d, a, 1.0, 20)
    K = po.Op(1, d, a, 1.0, 20, offside=off)
    assert rel(K.mul_padded(M), L.mul_padded(M)) < 1e-12


def test_negative_wind_is_reversed_stencil(rng):
    # UpwindDifference with coefficient -1: backward window, reverse(w)*(-1)^d
    for d, a in [(1, 1), (1, 2), (2, 2), (2, 3)]:
        n = 12
        pos = po.Op(1, d, a, 1.0, n, coeff=1.0)
        neg = po.Op(1, d, a, 1.0, n, coeff=-1.0)
        SL = pos.SL
        x = rng.random(n + 2)
        out = neg.mul_padded(x)
        sgn = (-1) ** d
        for i in range(pos.BPC + 1, n - pos.BPC + 1):   # interior rows (1-based)
            win = x[i - SL + 1: i + 1]  # padded indices i-SL+2 .. i+1
            exp = -1.0 * np.dot(sgn * pos.stencil[::-1], win)
            assert math.isclose(out[i - 1], exp, rel_tol=1e-12, abs_tol=1e-12)


# ---------------------------------------------------------------------------------------
# the N-D drivers

def test_axis_lines_match_1d(rng):
    # differentiation_dimension.jl: mul! along axis N equals the 1-D operator on each line
    M = rng.random((9, 22))  # Julia size (22, 9)
    op = po.Op(0, 4, 4, 0.1, 20)
    out = po.mul_nd(op, M, (9, 20))
    for j in range(9):
        assert rel(out[j], op.mul_padded(M[j])) < 1e-15


def test_ghost_apply_sums_single_ops(rng):
    u = rng.standard_normal((14, 17))    # Julia (17, 14)
    bcx = po.robin((1.0, 0.5, 0.2), (1.0, 0.5, -0.3), 0.1)
    bcy = po.dirichlet0()
    Lx = po.Op(0, 2, 4, 0.1, 17, axis=0)
    Ly = po.Op(1, 1, 2, 0.1, 14, axis=1, coeff=rng.standard_normal(14))
    both = po.ghost_apply([Lx, Ly], [bcx, bcy], u)
    sep = po.ghost_apply([Lx], [bcx, bcy], u) + po.ghost_apply([Ly], [bcx, bcy], u)
    assert rel(both, sep) < 1e-15
    # a line along x equals the 1-D operator on [ghost; line; ghost]
    for j in range(14):
        gl, gr = bcx.pad(u[j])
        exp = Lx.mul_padded(np.concatenate([[gl], u[j], [gr]]))
        assert rel(po.ghost_apply([Lx], [bcx, bcy], u)[j], exp) < 1e-15


def test_composite_padded_on_separable_field():
    # 2D_3D_fast_multiplication.jl:5-24: on cos(0.1i)+sin(0.1j) the composite result
    # equals the aligned sum of the single-axis results
    N = 40
    i = np.arange(1, N + 3)[:, None]
    j = np.arange(1, N + 3)[None, :]
    Mj = np.cos(0.1 * i) + np.sin(0.1 * j)     # Julia M[i, j]
    M = np.ascontiguousarray(Mj.T)            # numpy (j, i)
    Lxx = po.Op(0, 2, 2, 1.0, N, axis=0)
    Lyy = po.Op(0, 2, 2, 1.0, N, axis=1)
    comp = po.composite_mul_padded([Lxx, Lyy], M)
    ax = po.mul_nd(Lxx, M, (N + 2, N))[1:N + 1, :]
    ay = po.mul_nd(Lyy, M, (N, N + 2))[:, 1:N + 1]
    assert rel(comp, ax + ay) < 1e-12


def test_composite_padded_transverse_view_quirk(rng):
    # the composite else-branch views M[1:n1+2, 1:n2] (derivative_operator_functions.jl:
    # 429-460): on a non-separable field it is the transverse-shifted sum
    N = 12
    M = rng.random((N + 2, N + 2))
    Lxx = po.Op(0, 2, 2, 1.0, N, axis=0)
    Lyy = po.Op(0, 2, 2, 1.0, N, axis=1)
    comp = po.composite_mul_padded([Lxx, Lyy], M)
    ax = po.mul_nd(Lxx, M, (N + 2, N))[0:N, :]    # rows 1:n (not 2:n+1)
    ay = po.mul_nd(Lyy, M, (N, N + 2))[:, 0:N]
    assert rel(comp, ax + ay) < 1e-12
    aligned = po.mul_nd(Lxx, M, (N + 2, N))[1:N + 1, :] + po.mul_nd(Lyy, M, (N, N + 2))[:, 1:N + 1]
    assert rel(comp, aligned) > 1e-3


def test_oracle_rejects_short_lines():
    op = po.Op(0, 2, 6, 1.0, 4)    # BPC = 2 needs len >= 6
    with pytest.raises(ValueError):
        op.mul_padded(np.zeros(6))


def test_planted_mistake_is_caught(fixtures):
    # A plausible slip -- the right-boundary stencil not reflected -- must fail the
    # golden matrix check.
    e = fixtures["centered_uniform"][0]
    op = po.Op(0, e["d"], e["a"], e["dx"], e["N"])
    op.high = op.low.copy()
    assert rel(op.matrix(), np.array(e["A"])) > 1e-3


# ---------------------------------------------------------------------------------------
# PeriodicBC

def test_periodic_second_order_check():
    # 2nd_order_check.jl:3-38: L2*Q*u0 equals the periodic tridiagonal (1,-2,1)/h^2, and
    # L1*Q*(L1*Q*u0) the periodic (1/4, 0, -1/2, 0, 1/4)/h^2 stencil
    ordr = 2
    h = 0.025 * math.pi
    N = int(round(2 * (math.pi / h)))
    x = -math.pi + h * np.arange(N)
    u0 = np.cos(x)
    L1 = po.Op(0, 1, ordr, h, N)
    L2 = po.Op(0, 2, ordr, h, N)
    Q = po.periodic()
    M = np.diag(np.full(N, -2.0)) + np.diag(np.ones(N - 1), 1) + np.diag(np.ones(N - 1), -1)
    M[-1, 0] = M[0, -1] = 1.0
    M /= h * h
    got2 = po.ghost_apply([L2], [Q], u0, bpv_1d=True)
    assert rel(got2, M @ u0) < 1e-12
    A = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            if i == j:
                A[i, j] = -0.5
            if abs(i - j) == 2:
                A[i, j] = 0.25
    A[-2, 0] = A[-1, 1] = A[0, -2] = A[1, -1] = 0.25
    A /= h * h
    t = po.ghost_apply([L1], [Q], u0, bpv_1d=True)
    assert rel(po.ghost_apply([L1], [Q], t, bpv_1d=True), A @ u0) < 1e-12
    assert np.max(-np.cos(x) - got2) < 0.001


def test_periodic_upwind_bpv(fixtures):
    # upwind_operators_interface.jl:177-566: L * (PeriodicBC * x[1:N]) == analyticL * bpv
    rng = np.random.default_rng(4)
    for e in fixtures["upwind_uniform"]:
        n = len(e["A"])
        op = po.Op(1, e["d"], e["a"], e["dx"], n, coeff=e["coeff"])
        x = rng.random(n)
        bpv = np.concatenate([[x[-1]], x, [x[0]]])
        got = po.ghost_apply([op], [po.periodic()], x, bpv_1d=True)
        assert rel(got, np.array(e["A"]) @ bpv) < 1e-13, e["name"]


def test_periodic_nd_padding_is_edge_planes(rng):
    # multi_dim_bc_operators.jl:221-252: an all-periodic N-D BC pads with the first and last
    # planes of u; a line wrap is what the atomic BC gives a vector (bc_operators.jl:192)
    u = rng.standard_normal((7, 12))   # Julia (12, 7)
    L = po.Op(0, 2, 2, 0.1, 12, axis=0)
    nd = po.ghost_apply([L], [po.periodic_nd(), po.periodic_nd()], u)
    wrap = po.ghost_apply([L], [po.periodic(), po.periodic()], u)
    for j in range(7):
        line = u[j]
        assert rel(nd[j], L.mul_padded(np.concatenate([[line[0]], line, [line[-1]]]))) < 1e-15
        assert rel(wrap[j], L.mul_padded(np.concatenate([[line[-1]], line, [line[0]]]))) < 1e-15


# ---------------------------------------------------------------------------------------
# GeneralBC

def test_general_bc_equals_robin(fixtures):
    # robin.jl:92-110: GeneralBC([-10, 1, 6], [-10, 1, 6], dx, 3) pads u = 1:10 with the same
    # ghosts as RobinBC((1, 6, 10), (1, 6, 10), dx, 3): -4/10 and 125/12, uniform and vector dx
    u = np.arange(1.0, 11.0)
    for dx in (1.0, np.ones(10)):
        g = po.general([-10.0, 1.0, 6.0], [-10.0, 1.0, 6.0], dx, 3)
        gl, gr = g.pad(u)
        assert math.isclose(gl, -4 / 10, rel_tol=1e-12)
        assert math.isclose(gr, 125 / 12, rel_tol=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dx", [0.1, "vector"])
@pytest.mark.parametrize("nl", [3, 4, 5])
def test_general_bc_polynomial_ghost(order, dx, nl):
    # alpha0 + alpha1 g + alpha2 g' + alpha3 g'' + ... = 0 at the ghost point: the i-th
    # derivative formula on points 1..order+i is exact to degree order+i-1, so for a
    # polynomial p of degree order, choosing alpha0 from p makes the left ghost equal p at
    # the ghost point (x0 = 0, u at x = h, 2h, ...)
    rng = np.random.default_rng(order * 10 + nl)
    n = 12
    h = 0.1
    xs = h * np.arange(1, n + 1)
    p = np.polynomial.Polynomial(rng.standard_normal(order + 1))
    coef = [0.7, -1.3, 0.4, 0.9][: nl - 1]
    alpha0 = -sum(cf * p.deriv(k)(0.0) if k else cf * p(0.0) for k, cf in enumerate(coef))
    alpha = [alpha0] + coef
    dxa = h if dx == 0.1 else np.full(n + 1, h)
    g = po.general(alpha, alpha, dxa, order)
    gl, _ = g.pad(p(xs))
    assert math.isclose(gl, p(0.0), rel_tol=1e-8, abs_tol=1e-8)


def test_general_bc_planted_mistake_is_caught():
    # a wrong sign on b_l must break the Robin equivalence
    g = po.general([-10.0, 1.0, 6.0], [-10.0, 1.0, 6.0], 1.0, 3)
    g.b_l = -g.b_l
    gl, _ = g.pad(np.arange(1.0, 11.0))
    assert not math.isclose(gl, -4 / 10, rel_tol=1e-6)
