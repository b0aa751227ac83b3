"""Pins the oracle's adaptive Runge-Kutta integrator (oracle/pyoracle.py: tsit5), the checker of
the device's dop_tsit5: the Tsitouras 5(4) tableau against the Runge-Kutta order conditions
(order 5 for b, exactly order 4 for the embedded bhat = b - btilde, c = row sums of A), and the
integrator against the exact solution of linear method-of-lines systems (matrix exponential)."""
import numpy as np
import pytest
import scipy.linalg

import pyoracle as po


def _conds(w, A, c):
    Ac = A @ c
    return [
        (w.sum(), 1.0), (w @ c, 1 / 2), (w @ c ** 2, 1 / 3), (w @ Ac, 1 / 6),
        (w @ c ** 3, 1 / 4), (w @ (c * Ac), 1 / 8), (w @ (A @ c ** 2), 1 / 12), (w @ (A @ Ac), 1 / 24),
        (w @ c ** 4, 1 / 5), (w @ (c ** 2 * Ac), 1 / 10), (w @ (c * (A @ c ** 2)), 1 / 15),
        (w @ (c * (A @ Ac)), 1 / 30), (w @ Ac ** 2, 1 / 20), (w @ (A @ c ** 3), 1 / 20),
        (w @ (A @ (c * Ac)), 1 / 40), (w @ (A @ (A @ c ** 2)), 1 / 60), (w @ (A @ (A @ Ac)), 1 / 120),
    ]


def test_tsit5_tableau_order_conditions():
    A = np.zeros((7, 7))
    A[:, :6] = po.TSIT5_A
    c = A.sum(axis=1)
    assert np.allclose(c, po.TSIT5_C, rtol=0, atol=1e-15)
    b = po.TSIT5_B
    assert np.array_equal(b[:6], A[6, :6]) and b[6] == 0.0     # first same as last
    for got, want in _conds(b, A, c):                           # order 5
        assert abs(got - want) < 1e-13
    bhat = b - po.TSIT5_BTILDE
    conds = _conds(bhat, A, c)
    for got, want in conds[:8]:                                 # order 4
        assert abs(got - want) < 1e-13
    assert max(abs(g - w) for g, w in conds[8:]) > 1e-5        # and not order 5


def _heat(n, kind):
    """du/dt = A u + g: the second-difference matrix with Dirichlet (g from the boundary
    values) or homogeneous Neumann-like Robin ends, on (0, 1)."""
    h = 1.0 / (n + 1)
    A = (np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)) / h ** 2
    g = np.zeros(n)
    if kind == "dirichlet":
        g[0], g[-1] = 1.0 / h ** 2, -0.5 / h ** 2
    else:
        A[0, 0] = A[-1, -1] = -1.0 / h ** 2
        A -= 0.3 * np.eye(n)
    return A, g


@pytest.mark.parametrize("kind", ["dirichlet", "robin"])
@pytest.mark.parametrize("rtol", [1e-3, 1e-6])
def test_tsit5_matches_matrix_exponential(kind, rtol):
    n = 24
    A, g = _heat(n, kind)
    x = np.linspace(0, 1, n + 2)[1:-1]
    u0 = np.sin(3 * np.pi * x) + x
    T = 0.05
    f = lambda u: A @ u + g     # noqa: E731
    u, t, na, nr = po.tsit5(f, u0, 0.0, T, atol=rtol * 1e-3, rtol=rtol)
    assert t == pytest.approx(T, abs=1e-14)
    w = np.linalg.solve(A, g)
    exact = scipy.linalg.expm(T * A) @ (u0 + w) - w
    err = np.abs(u - exact).max() / np.abs(exact).max()
    # the local-error control keeps the global error within a small multiple of rtol
    assert err < 20 * rtol
    assert na > 0


def test_tsit5_planted_mistake_is_caught():
    # a wrong weight (b3 off by 1%: the update is no longer consistent) breaks the agreement
    # with the exact solution
    n = 24
    A, g = _heat(n, "dirichlet")
    u0 = np.linspace(0, 1, n)
    f = lambda u: A @ u + g     # noqa: E731
    w = np.linalg.solve(A, g)
    exact = scipy.linalg.expm(0.05 * A) @ (u0 + w) - w
    keep = po.TSIT5_A[6, 2]
    try:
        po.TSIT5_A[6, 2] = keep * 1.01
        u, _, _, _ = po.tsit5(f, u0, 0.0, 0.05, atol=1e-9, rtol=1e-6)
    finally:
        po.TSIT5_A[6, 2] = keep
    assert np.abs(u - exact).max() / np.abs(exact).max() > 1e-4
