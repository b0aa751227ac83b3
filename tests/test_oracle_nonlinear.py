"""Pins the oracle's nonlinear_diffusion (oracle/pyoracle.py) against the reference's
Fast_Diffusion.jl setup (d/dx(k du/dx) with k = u^-2 of u = 1/sqrt(x^2 + 1): -(x^2 + 1)^(-3/2)),
against what the Leibniz rule fixes (p = 1 leaves D^{f+s} q), and a planted mistake."""
import numpy as np

import pyoracle as po


def test_fast_diffusion_rhs_at_t0():
    n = 100
    h = 1.0 / (n + 1)
    x = h * np.arange(0, n + 2)                    # padded grid: ghosts at x = 0 and x = 1
    q = 1.0 / np.sqrt(x ** 2 + 1.0)                # Dirichlet ghosts = the boundary values
    p = q ** -2
    du = po.nonlinear_diffusion(1, 1, 2, p, q, h, n)
    exact = -(x[1:-1] ** 2 + 1.0) ** -1.5
    assert np.abs(du - exact).max() < 1e-3


def test_constant_coefficient_reduces_to_one_operator(rng):
    n = 40
    h = 0.05
    q = rng.standard_normal(n + 2)
    for s, f, a in ((1, 1, 2), (2, 1, 4), (3, 0, 2), (2, 2, 4)):
        du = po.nonlinear_diffusion(s, f, a, np.ones(n + 2), q, h, n)
        ref = po.Op(0, f + s, a, h, n).mul_padded(q)
        assert np.abs(du - ref).max() <= 1e-9 * np.abs(ref).max()


def test_planted_mistake_is_caught(rng):
    n = 100
    h = 1.0 / (n + 1)
    x = h * np.arange(0, n + 2)
    q = 1.0 / np.sqrt(x ** 2 + 1.0)
    p = q ** -2
    du = po.nonlinear_diffusion(1, 1, 2, q, p, h, n)    # p and q swapped
    assert np.abs(du + (x[1:-1] ** 2 + 1.0) ** -1.5).max() > 1e-2
