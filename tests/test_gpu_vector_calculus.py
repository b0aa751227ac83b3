"""GPU parity of Gradient / Divergence / Curl (libdop padded plans with the transverse view of
vector_calculus_convolutions.jl) against the oracle (pyoracle.gradient / divergence / curl) on
random non-separable fields: FP64 1e-12, FP32 storage 1e-5 (normwise); and the reference tests'
analytic checks (gradient_operator.jl, divergence_operator.jl, curl_operator.jl) on the device."""
import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}


def rel(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300)


def oops(A):
    out = []
    for L in A.ops:
        o = po.Op(0, 1, L.approximation_order, L.dx, L.len, axis=L.N - 1)
        o.coeff = L.coefficients.copy()
        o.coeff_kind = L.coeff_kind
        out.append(o)
    return out


def _dx(rng, n, uniform):
    return 0.1 if uniform else 0.05 * (1 + 0.5 * rng.random(n + 1))


@pytest.mark.parametrize("shape", [(30,), (17, 23), (9, 11, 13)])
@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gradient_and_divergence_match_oracle(rng, shape, uniform, dtype):
    jsize = tuple(reversed(shape))
    nd = len(shape)
    dx = tuple(_dx(rng, n, uniform) for n in jsize)
    G = D.Gradient(4, dx, jsize, coeff_func=rng.uniform(0.5, 1.5, jsize[0]) if nd == 1 else 1)
    Dv = 1.7 * D.Divergence(2, dx, jsize)
    print(f"[vc] {shape} uniform={uniform} {dtype}", flush=True)
    M = rng.uniform(-1, 1, tuple(s + 2 for s in shape))
    Mq = torch.from_numpy(M).to("cuda", dtype)
    got = (G * Mq).double().cpu().numpy()
    ref = po.gradient(oops(G), Mq.double().cpu().numpy())
    assert got.shape == ref.shape
    assert rel(got, ref) <= TOL[dtype]
    V = rng.uniform(-1, 1, (nd,) + tuple(s + 2 for s in shape))
    Vq = torch.from_numpy(V).to("cuda", dtype)
    got = (Dv * Vq).double().cpu().numpy()
    ref = po.divergence(oops(Dv), Vq.double().cpu().numpy())
    assert got.shape == ref.shape
    assert rel(got, ref) <= TOL[dtype]


@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_curl_matches_oracle(rng, uniform, dtype):
    n = 12
    dx = tuple(_dx(rng, n, uniform) for _ in range(3))
    C = D.Curl(4, dx, (n, n, n))
    V = rng.uniform(-1, 1, (3, n + 2, n + 2, n + 2))
    Vq = torch.from_numpy(V).to("cuda", dtype)
    got = (C * Vq).double().cpu().numpy()
    ref = po.curl(oops(C), Vq.double().cpu().numpy())
    assert got.shape == ref.shape == (3, n, n, n)
    assert rel(got, ref) <= TOL[dtype]


def test_reference_analytic_checks_on_device():
    x = np.arange(-5.0, 5.0 + 1e-12, 1.25)
    n = len(x) - 2
    X, Y, Z = [np.ascontiguousarray(np.transpose(m)) for m in np.meshgrid(x, x, x, indexing="ij")]
    s = (slice(1, -1),) * 3
    for dx in ((1.25,) * 3, tuple(1.25 * np.ones(len(x) - 1) for _ in range(3))):
        g = D.Gradient(4, dx, (n, n, n)) * torch.from_numpy(X ** 2 + Y ** 2 + Z ** 2).cuda()
        g = g.cpu().numpy()
        for k, c in enumerate((X, Y, Z)):
            assert np.allclose(g[k], 2 * c[s], atol=1e-3, rtol=0)
        u = np.stack([X ** 2 + Y ** 2, Y ** 2 + Z ** 2, X ** 2 + Z ** 2])
        d = (D.Divergence(4, dx, (n, n, n)) * torch.from_numpy(u).cuda()).cpu().numpy()
        assert np.allclose(d, 2 * (X + Y + Z)[s], atol=1e-3, rtol=0)
        u = np.stack([Y ** 2 + Z ** 2, X ** 2 + Z ** 2, X ** 2 + Y ** 2])
        A = D.Curl(4, dx, (n, n, n))
        c = (A * torch.from_numpy(u).cuda()).cpu().numpy()
        ana = np.stack([2 * (Y - Z)[s], 2 * (Z - X)[s], 2 * (X - Y)[s]])
        assert np.allclose(c, ana, atol=1e-3, rtol=0)
        c2 = ((2 * A) * torch.from_numpy(u).cuda()).cpu().numpy()
        assert np.allclose(c2, 2 * c, atol=1e-3, rtol=0)


def test_vector_calculus_errors():
    with pytest.raises(AssertionError):
        D.Curl(2, (0.1, 0.1, 0.1), (5, 6, 5))
    G = D.Gradient(2, (0.1, 0.1), (5, 6))
    with pytest.raises(AssertionError):
        G * torch.zeros((8, 7, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        G * torch.zeros((8, 7), dtype=torch.float64)
