"""GPU parity of nonlinear_diffusion! (dop_nonlinear_diffusion) against the oracle's
restatement (derivative_operator.jl:34-66): FP64 1e-12, FP32 storage 1e-5 (normwise), on
random p, q, uniform and non-uniform dx, several (second, first, approximation) orders, and
the Fast_Diffusion.jl right-hand side on the device."""
import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("s,f,a", [(1, 1, 2), (2, 1, 4), (3, 0, 2), (2, 2, 6), (1, 0, 4)])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nonlinear_diffusion_matches_oracle(rng, s, f, a, uniform, dtype):
    n = 57
    dx = 0.05 if uniform else 0.05 * (1 + 0.5 * rng.random(n + 1))
    p = torch.from_numpy(rng.uniform(0.5, 1.5, n + 2)).to("cuda", dtype)
    q = torch.from_numpy(rng.standard_normal(n + 2)).to("cuda", dtype)
    got = D.nonlinear_diffusion(s, f, a, p, q, dx, n).double().cpu().numpy()
    ref = po.nonlinear_diffusion(s, f, a, p.double().cpu().numpy(), q.double().cpu().numpy(), dx, n)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_fast_diffusion_rhs_on_device():
    n = 100
    h = 1.0 / (n + 1)
    u = torch.from_numpy(1.0 / np.sqrt((h * np.arange(1, n + 1)) ** 2 + 1.0)).cuda()
    bc = D.DirichletBC(1.0, 2.0 ** -0.5)
    q = (bc * u).materialize()
    p = q ** -2
    du = torch.empty(n, dtype=torch.float64, device="cuda")
    D.nonlinear_diffusion_(du, 1, 1, 2, p, q, h, n)
    x = h * np.arange(1, n + 1)
    assert np.abs(du.cpu().numpy() + (x ** 2 + 1.0) ** -1.5).max() < 1e-3


def test_nonlinear_diffusion_errors():
    q = torch.ones(12, dtype=torch.float64, device="cuda")
    with pytest.raises(AssertionError):
        D.nonlinear_diffusion(1, 1, 1, q, q, 0.1, 10)
    with pytest.raises(AssertionError):
        D.nonlinear_diffusion(1, 1, 2, q, q, 0.1, 11)
    with pytest.raises(TypeError):
        D.nonlinear_diffusion(1, 1, 2, q.cpu(), q.cpu(), 0.1, 10)


@pytest.mark.parametrize("s,f,a", [(1, 1, 2), (2, 1, 4), (3, 0, 2), (2, 2, 6), (1, 0, 4), (4, 1, 2)])
@pytest.mark.parametrize("n", [61, 4096, 4097])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nonlinear_diffusion_four_point_form_bitwise(rng, s, f, a, n, dtype):
    """Uniform grids take the four-point kernel (windows read once per group of points); it
    must equal the per-point form bit for bit, at the ends and for lengths not a multiple of
    four included."""
    import os
    from diffeqoperators_amd import nonlinear as _nonlinear
    print(f"case: nld4 s={s} f={f} a={a} n={n} {dtype}", flush=True)
    dx = 1.0 / (n + 1)
    p = torch.from_numpy(rng.uniform(0.5, 1.5, n + 2)).to("cuda", dtype)
    q = torch.from_numpy(rng.standard_normal(n + 2)).to("cuda", dtype)
    outs = []
    for generic in (False, True):
        if generic:
            os.environ["DOP_FORCE_GENERIC"] = "1"
        _nonlinear._CACHE.clear()
        try:
            outs.append(D.nonlinear_diffusion(s, f, a, p, q, dx, n))
        finally:
            os.environ.pop("DOP_FORCE_GENERIC", None)
            _nonlinear._CACHE.clear()
    assert torch.equal(outs[0], outs[1])
    ref = po.nonlinear_diffusion(s, f, a, p.double().cpu().numpy(), q.double().cpu().numpy(), dx, n)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    got = outs[0].double().cpu().numpy()
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
