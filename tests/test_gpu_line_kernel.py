"""GPU: the single-axis kernel (plans whose operators all act on one axis: 1-D problems, one
direction of an N-D array, the padded plans of Gradient / Divergence / Curl and
nonlinear_diffusion!) against the generic per-point kernel, which it must match bit for bit
(the same arithmetic per point), in every storage / accumulation mode and with accumulate."""
import os

import numpy as np
import pytest
import torch

import diffeqoperators_amd as D
from diffeqoperators_amd import nonlinear as _nonlinear

pytestmark = pytest.mark.gpu


def _fresh_plan(A, u, generic):
    if generic:
        os.environ["DOP_FORCE_GENERIC"] = "1"
    try:
        A2 = D.DiffEqOperatorCombination(A.ops) if isinstance(A, D.DiffEqOperatorCombination) else A
        A2._plans = {}
        return D.plans_for(A2, u)[0]
    finally:
        os.environ.pop("DOP_FORCE_GENERIC", None)


def _check(A, u):
    a = torch.empty_like(u)
    b = torch.empty_like(u)
    _fresh_plan(A, u, False).apply(a, u)
    _fresh_plan(A, u, True).apply(b, u)
    assert torch.equal(a, b)
    a2, b2 = a.clone(), b.clone()
    _fresh_plan(A, u, False).apply(a2, u, accumulate=True)
    _fresh_plan(A, u, True).apply(b2, u, accumulate=True)
    assert torch.equal(a2, b2)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("d,approx", [(1, 2), (2, 4), (3, 3), (4, 6), (2, 8)])
@pytest.mark.parametrize("uniform", [True, False])
def test_line_kernel_1d_bitwise(rng, dtype, d, approx, uniform):
    n = 3001
    print(f"case: line1d d={d} a={approx} uniform={uniform} {dtype}", flush=True)
    dx = 1.0 / (n + 1) if uniform else (1 + 0.5 * rng.random(n + 1)) / (n + 1)
    L = D.CenteredDifference(d, approx, dx, n) + D.UpwindDifference(1, 2, dx, n, rng.standard_normal(n))
    A = L * D.RobinBC((1.0, 0.5, 0.3), (1.0, 0.4, -0.2), 1.0 / (n + 1), 2)
    u = torch.from_numpy(rng.uniform(-1, 1, n)).to("cuda").to(dtype)
    _check(A, u)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_line_kernel_one_axis_of_3d_bitwise(rng, dtype):
    shape = (20, 33, 41)
    jsize = tuple(reversed(shape))
    print(f"case: line3d axis 2 {dtype}", flush=True)
    L = D.CenteredDifference[2](2, 4, 0.05 * (1 + 0.5 * rng.random(jsize[1] + 1)), jsize[1])
    A = L * D.compose(*D.Dirichlet0BC(float, jsize))
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    _check(A, u)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_padded_plans_bitwise(rng, dtype):
    n = 24
    print(f"case: padded vector calculus / nonlinear diffusion {dtype}", flush=True)
    dx = tuple(0.1 * (1 + 0.5 * rng.random(n + 1)) for _ in range(3))
    V = torch.from_numpy(rng.uniform(-1, 1, (3, n + 2, n + 2, n + 2))).to("cuda").to(dtype)
    M = V[0].contiguous()
    outs = []
    for generic in (False, True):
        if generic:
            os.environ["DOP_FORCE_GENERIC"] = "1"
        _nonlinear._CACHE.clear()   # its plans are cached per (orders, grid, dtype)
        try:
            g = D.Gradient(4, dx, (n, n, n)) * M
            dv = D.Divergence(2, dx, (n, n, n)) * V
            c = D.Curl(4, dx, (n, n, n)) * V
            m = 4097
            p = torch.from_numpy(rng.uniform(0.5, 1.5, m + 2)).to("cuda").to(dtype) if not generic else outs[0][3]
            q = torch.from_numpy(rng.uniform(-1, 1, m + 2)).to("cuda").to(dtype) if not generic else outs[0][4]
            nl = D.nonlinear_diffusion(2, 1, 4, p, q, 1.0 / (m + 1), m)
        finally:
            os.environ.pop("DOP_FORCE_GENERIC", None)
            _nonlinear._CACHE.clear()
        outs.append((g, dv, c, p, q, nl))
    for x, y in zip(outs[0], outs[1]):
        assert torch.equal(x, y)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n", [4, 8, 12, 16, 20, 3000])
@pytest.mark.parametrize("d,approx", [(1, 2), (2, 4), (4, 6), (2, 8)])
@pytest.mark.parametrize("uniform", [True, False])
def test_line_kernel_groups_of_four_bitwise(rng, dtype, n, d, approx, uniform):
    """Rows of a multiple of four points take the four-point form: boundary rows at both ends,
    interior rows inside the edge groups, rows whose interior holds no whole group."""
    if d + approx > n:
        pytest.skip("the boundary stencil is longer than the grid")
    print(f"case: line4 n={n} d={d} a={approx} uniform={uniform} {dtype}", flush=True)
    dx = 1.0 / (n + 1) if uniform else (1 + 0.5 * rng.random(n + 1)) / (n + 1)
    L = D.CenteredDifference(d, approx, dx, n)
    A = L * D.RobinBC((1.0, 0.5, 0.3), (1.0, 0.4, -0.2), 1.0 / (n + 1), 2)
    u = torch.from_numpy(rng.uniform(-1, 1, n)).to("cuda").to(dtype)
    _check(A, u)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("axis", [1, 2, 3])
def test_line_kernel_groups_of_four_3d_bitwise(rng, dtype, axis):
    """A 3-D array with 44 points per row: the active axis is the row axis (windows read in
    groups of four) or a strided axis (four points share one tap row)."""
    shape = (12, 9, 44)
    jsize = tuple(reversed(shape))
    m = jsize[axis - 1]
    print(f"case: line4 3d axis {axis} {dtype}", flush=True)
    L = D.CenteredDifference[axis](2, 4, 0.05 * (1 + 0.5 * rng.random(m + 1)), m)
    A = L * D.compose(*D.Dirichlet0BC(float, jsize))
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    _check(A, u)
