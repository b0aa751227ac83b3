"""GPU tests of the device-resident adaptive integrator (dop_tsit5_*) through the C ABI.

Parity: the device run and the oracle's restatement (pyoracle.tsit5 with the oracle's FP64
apply as the right-hand side) take the same accepted / rejected steps and end within 1e-9
(relative, FP64 storage); the reference's own heat-equation checks (heat_eqn.jl:4-60) hold at
t = 0.1, 0.2, ..., 1.0.  OrdinaryDiffEq's own step sequence is parity unpinned (an absent,
unpinned dependency)."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    po.set_threads(min(16, os.cpu_count() or 1))
    yield
    po.set_threads(1)


def oop(L):
    o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
              axis=L.N - 1, offside=L.offside)
    o.coeff = L.coefficients.copy()
    o.coeff_kind = L.coeff_kind
    return o


def obc(bc):
    return po.BC(bc.a_l, bc.b_l, bc.a_r, bc.b_r)


def test_heat_dirichlet_matches_oracle_and_reference_checks():
    # heat_eqn.jl:4-28: x = -pi:2pi/511:pi, u0 = -(x - 0.5)^2 + 1/12, Dirichlet from u_analytic
    x = np.linspace(-np.pi, np.pi, 512)
    dx = 2 * np.pi / 511
    ua = lambda z: -(z - 0.5) ** 2 + 1 / 12     # noqa: E731
    u0 = ua(x)
    A = D.CenteredDifference(2, 2, dx, 512)
    bc = D.DirichletBC(ua(-np.pi - dx), ua(np.pi + dx))
    u = torch.from_numpy(u0.copy()).cuda()
    print("[tsit5] heat dirichlet 512", flush=True)
    I = D.Tsit5(A * bc, u, (0.0, 1.0))
    for k in range(1, 11):
        I.advance_to(0.1 * k)
        v = u.cpu().numpy()
        # the reference's assertions: soln(t)[1] ≈ u0[1], soln(t)[end] ≈ u0[end] (rtol 0.05)
        assert v[0] == pytest.approx(u0[0], rel=0.05)
        assert v[-1] == pytest.approx(u0[-1], rel=0.05)
    print(f"[tsit5] device: {I.status()}", flush=True)
    # one segment, (0, 0.1), on the device and in the oracle's restatement
    f = lambda w: po.ghost_apply([oop(A)], [obc(bc)], w, bpv_1d=True)   # noqa: E731
    w, t, na, nr = po.tsit5(f, u0.copy(), 0.0, 0.1)
    J = D.Tsit5(A * bc, torch.from_numpy(u0.copy()).cuda(), (0.0, 0.1))
    J.solve()
    sj = J.status()
    print(f"[tsit5] segment: device {sj}, oracle na={na} nr={nr}", flush=True)
    assert sj["t"] == pytest.approx(t, abs=1e-14)
    assert (sj["naccept"], sj["nreject"]) == (na, nr)
    got = J.u.cpu().numpy()
    assert np.abs(got - w).max() <= 1e-9 * np.abs(w).max()


def test_heat_neumann_reference_checks():
    # heat_eqn.jl:31-52: the boundary derivatives stay at their Neumann values (atol 0.1)
    N = 128
    dx = 2 * np.pi / (N - 1)
    x = np.arange(-np.pi, np.pi + dx / 2, dx)[:N]
    u0 = -(x - 0.5) ** 2 + 1 / 12
    B = D.CenteredDifference(1, 2, dx, N - 2)
    Bu = (B * torch.from_numpy(u0).cuda()).cpu().numpy()
    ds, de = Bu[0], Bu[-1]
    A = D.CenteredDifference(2, 2, dx, N)
    bc = D.NeumannBC((ds, de), dx, 1)
    u = torch.from_numpy(u0.copy()).cuda()
    I = D.Tsit5(A * bc, u, (0.0, 1.0))
    cs = np.array([-11 / 6, 3.0, -3 / 2, 1 / 3]) / dx
    ce = -cs[::-1]
    for k in range(1, 11):
        I.advance_to(0.1 * k)
        v = u.cpu().numpy()
        assert np.sum(cs * v[:4]) == pytest.approx(ds, abs=1e-1)
        assert np.sum(ce * v[-4:]) == pytest.approx(de, abs=1e-1)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_2d_advection_diffusion_matches_oracle(rng, dtype):
    n1, n2 = 48, 40
    h = 1.0 / (n1 + 1)
    vx = rng.standard_normal(n1)
    ops = [D.CenteredDifference[1](2, 4, h, n1), D.CenteredDifference[2](2, 4, h, n2),
           D.UpwindDifference[1](1, 3, h, n1, -vx)]
    bcs = [D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.1), h), D.Dirichlet0BC(float)]
    Q = D.compose(D.MultiDimBC[1](bcs[0], (n1, n2)), D.MultiDimBC[2](bcs[1], (n1, n2)))
    G = (ops[0] + ops[1] + ops[2]) * Q
    u0 = rng.uniform(-1, 1, (n2, n1))
    scale = 0.05
    u = torch.from_numpy(u0.copy()).to("cuda", dtype)
    print(f"[tsit5] 2-D {dtype}", flush=True)
    I = D.Tsit5(G, u, (0.0, 0.02), scale=scale)
    I.solve()
    st = I.status()
    f = lambda w: scale * po.ghost_apply([oop(L) for L in ops], [obc(b) for b in bcs], w)   # noqa: E731
    w0 = u0 if dtype == torch.float64 else u0.astype(np.float32).astype(np.float64)
    w, t, na, nr = po.tsit5(f, w0, 0.0, 0.02)
    print(f"[tsit5] device {st}, oracle na={na} nr={nr}", flush=True)
    got = u.double().cpu().numpy()
    if dtype == torch.float64:
        assert (st["naccept"], st["nreject"]) == (na, nr)
        assert np.abs(got - w).max() <= 1e-9 * np.abs(w).max()
    else:
        # FP32 storage of the stages: the solution agrees to the integration tolerance
        assert abs(st["naccept"] - na) <= max(2, na // 20)
        assert np.abs(got - w).max() <= 1e-3 * np.abs(w).max()


def test_tsit5_errors():
    L = D.CenteredDifference(2, 2, 0.1, 10)
    with pytest.raises(TypeError):
        D.Tsit5(L, torch.zeros(12, dtype=torch.float64, device="cuda"), (0.0, 1.0))
    with pytest.raises(TypeError):
        D.Tsit5(L * D.Dirichlet0BC(float), torch.zeros(10, dtype=torch.float64), (0.0, 1.0))
