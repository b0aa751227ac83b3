"""GPU: persistent batched application and graph-captured RK4 over many independent
problems (uneven sizes, orders and BCs in one launch) against per-problem applies and the
FP64 oracle."""
import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu


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


def random_problem(rng, dtype):
    """1-D or 2-D; random d, a and BC (Dirichlet / Neumann / Robin)."""
    ndim = int(rng.integers(1, 3))
    d = int(rng.integers(1, 5))
    a = int(rng.choice([2, 4, 6, 8]))
    if ndim == 1:
        n = int(np.exp(rng.uniform(np.log(64), np.log(4096))))
        shape = (n,)
    else:
        n = int(np.exp(rng.uniform(np.log(64), np.log(300))))
        shape = (n, n)
    h = 1.0 / (n + 1)
    g = float(rng.uniform(-1, 1))
    kind = int(rng.integers(0, 3))
    if kind == 0:
        atom = D.DirichletBC(g, -g)
    elif kind == 1:
        atom = D.NeumannBC((g, 0.5 * g), h, 1)
    else:
        atom = D.RobinBC((1.0, 0.5, g), (1.0, 0.5, -g), h, 1)
    if ndim == 1:
        L = D.CenteredDifference(d, a, h, n)
        A = L * atom
        ops, bcs = [L], [atom]
    else:
        Lx = D.CenteredDifference[1](d, a, h, n)
        Ly = D.CenteredDifference[2](d, a, h, n)
        Q = D.compose(*D.MultiDimBC(atom, (n, n)))
        A = (Lx + Ly) * Q
        ops, bcs = [Lx, Ly], [atom, atom]
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    return A, u, ops, bcs, d, h


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_equals_single_applies(dtype):
    rng = np.random.default_rng(42)
    probs = [random_problem(rng, dtype) for _ in range(40)]
    ops = [p[0] for p in probs]
    us = [p[1] for p in probs]
    dus = [torch.full_like(u, float("nan")) for u in us]
    b = D.Batch(ops, us, dus)
    b.apply()
    torch.cuda.synchronize()
    for (A, u, _, _, _, _), du in zip(probs, dus):
        single = A * u
        assert torch.equal(du, single)


def test_batch_matches_oracle_f64():
    rng = np.random.default_rng(43)
    probs = [random_problem(rng, torch.float64) for _ in range(12)]
    dus = [torch.empty_like(p[1]) for p in probs]
    D.Batch([p[0] for p in probs], [p[1] for p in probs], dus).apply()
    for (A, u, ops, bcs, _, _), du in zip(probs, dus):
        ref = po.ghost_apply([oop(L) for L in ops], [po.BC(b.a_l, b.b_l, b.a_r, b.b_r) for b in bcs],
                             u.cpu().numpy())
        assert rel(du.cpu().numpy(), ref) < 1e-12


def _rk4_oracle(f, u, dt, nsteps):
    for _ in range(nsteps):
        k1 = f(u)
        k2 = f(u + 0.5 * dt * k1)
        k3 = f(u + 0.5 * dt * k2)
        k4 = f(u + dt * k3)
        u = u + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    return u


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_rk4_matches_oracle(dtype, tol):
    rng = np.random.default_rng(44)
    probs = [random_problem(rng, dtype) for _ in range(10)]
    scales, dts, u0s = [], [], []
    for A, u, ops, bcs, d, h in probs:
        s = -1.0 if d % 2 == 1 else float((-1) ** (d // 2 + 1))
        wsum = sum(float(np.abs(L.stencil_coefs).sum()) for L in ops)
        scales.append(s)
        dts.append(0.5 / wsum)
        u0s.append(u.to(torch.float64).cpu().numpy())
    rk = D.RK4([p[0] for p in probs], [p[1] for p in probs], scales, dts)
    nsteps = 5
    rk.step(nsteps)
    torch.cuda.synchronize()
    for (A, u, ops, bcs, d, h), s, dt, u0 in zip(probs, scales, dts, u0s):
        obcs = [po.BC(b.a_l, b.b_l, b.a_r, b.b_r) for b in bcs]
        f = lambda v: s * po.ghost_apply([oop(L) for L in ops], obcs, v)
        ref = _rk4_oracle(f, u0, dt, nsteps)
        assert rel(u.to(torch.float64).cpu().numpy(), ref) < tol


def test_rk4_graph_replays_are_consistent():
    # replaying the captured step k times equals k separate replays, with eager work and
    # synchronizations between them
    rng = np.random.default_rng(45)
    probs = [random_problem(rng, torch.float64) for _ in range(6)]
    us_a = [p[1].clone() for p in probs]
    us_b = [p[1].clone() for p in probs]
    ops = [p[0] for p in probs]
    sc = [-1.0] * len(probs)
    dts = [1e-7] * len(probs)
    ra = D.RK4(ops, us_a, sc, dts)
    rb = D.RK4(ops, us_b, sc, dts)
    ra.step(4)
    for _ in range(4):
        rb.step(1)
        torch.cuda.synchronize()
        _ = (us_b[0] * 2).sum().item()
    torch.cuda.synchronize()
    for x, y in zip(us_a, us_b):
        assert torch.equal(x, y)


def test_batch_and_rk4_reject_mismatched_buffers():
    # a buffer of another dtype or length than the plan's would be read or written out of
    # bounds on the device: refused before the batch is created
    n = 100
    h = 1.0 / (n + 1)
    A = D.CenteredDifference(2, 2, h, n) * D.Dirichlet0BC(float)
    u = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        D.Batch([A], [u], [torch.zeros(n - 1, dtype=torch.float64, device="cuda")])
    with pytest.raises(TypeError):
        D.Batch([A], [u], [torch.zeros(n, dtype=torch.float32, device="cuda")])
    with pytest.raises(ValueError):
        D.Batch([A], [u], [u])


def test_server_submit_rejects_mismatched_u0():
    n = 64
    h = 1.0 / (n + 1)
    A = (D.CenteredDifference[1](2, 4, h, n) + D.CenteredDifference[2](2, 4, h, n)) * \
        D.compose(*D.Dirichlet0BC(float, (n, n)))
    plan = D.plan_for_shape(A, (n, n), torch.float32)
    srv = D.Server(64 << 20)
    with pytest.raises(ValueError):
        srv.submit(plan, torch.zeros(n * n - 1, dtype=torch.float32).pin_memory(), 1.0, 1e-5, 2, 0.0)
    with pytest.raises(TypeError):
        srv.submit(plan, torch.zeros(n * n, dtype=torch.float64).pin_memory(), 1.0, 1e-5, 2, 0.0)
    with pytest.raises(ValueError):
        srv.submit(plan, torch.zeros(n * n, dtype=torch.float32), 1.0, 1e-5, 2, 0.0)
