"""GPU: method-of-lines right-hand sides that rebuild their boundary condition at every call.

The reference's callers build a new BC inside every RHS evaluation -- heat_equation.jl:38-43
(`bc = DirichletBC(cos(t*1.5), sin(t)); Δ*bc*u`) and KdV.jl:26-30 (a time-varying GeneralBC,
`mul!(du, A, bc*u)`).  Here the implicit `A*u` / `mul_` path keeps one plan per structure and
replaces its values in place (dop_plan_update).  Over 100 steps each: one plan is used, every
step's result equals a freshly created plan's bit for bit and the FP64 oracle's to 1e-12; a
descriptor of another structure is refused (DOP_EINVAL) and leaves the plan intact."""
import math
import time

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from diffeqoperators_amd import _abi
from diffeqoperators_amd.apply import make_desc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture
def plan_count(monkeypatch):
    """Number of plans created so far (counts dop_plan_create through the mirror)."""
    import diffeqoperators_amd.apply as ap
    made = []
    orig = ap.Plan.__init__

    def counting(self, *a, **k):
        made.append(1)
        orig(self, *a, **k)

    monkeypatch.setattr(ap.Plan, "__init__", counting)
    return lambda: len(made)


def test_heat_equation_time_dependent_dirichlet(plan_count):
    n = 100
    h = 1.0 / (n + 1)
    x = h * np.arange(1, n + 1)
    Δ = D.CenteredDifference(2, 2, h, n)
    u = torch.from_numpy(np.sin(2 * np.pi * x)).to("cuda")
    dt = 0.2 * h * h
    made = 0
    t_upd = []
    for k in range(100):
        t = k * dt
        bc = D.DirichletBC(math.cos(1.5 * t), math.sin(t))
        torch.cuda.synchronize()
        c0 = plan_count()
        t0 = time.perf_counter()
        du = Δ * bc * u
        torch.cuda.synchronize()
        t_upd.append(time.perf_counter() - t0)
        made += plan_count() - c0
        fresh = D.plans_for(Δ * D.DirichletBC(math.cos(1.5 * t), math.sin(t)), u)[0]
        ref_fresh = torch.empty_like(u)
        fresh.apply(ref_fresh, u)
        assert torch.equal(du, ref_fresh), k
        uh = u.cpu().numpy()
        ref = po.ghost_apply([po.Op(0, 2, 2, h, n)], [po.dirichlet(math.cos(1.5 * t), math.sin(t))], uh)
        assert rel(du.cpu().numpy(), ref) <= 1e-12, k
        u = u + dt * du
    assert made <= 1   # one plan for the structure (none if an earlier test made it)
    # per-call cost of a fresh plan (what every call paid before plans were shared)
    t_new = []
    for k in range(20):
        bc = D.DirichletBC(math.cos(1.5 * k), math.sin(k))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p = D.plans_for(Δ * bc, u)[0]
        out = torch.empty_like(u)
        p.apply(out, u)
        torch.cuda.synchronize()
        t_new.append(time.perf_counter() - t0)
    print(f"  heat RHS with a new BC: {np.median(t_upd) * 1e6:.0f} us per call with the shared "
          f"plan, {np.median(t_new) * 1e6:.0f} us with a new plan per call", flush=True)


def phi(x, t):
    return 0.5 / np.cosh((x - t) / 2) ** 2


@pytest.mark.parametrize("offside", [0, 1])
def test_kdv_time_dependent_general_bc(offside, plan_count):
    N = 21
    dx = 1.0 / (N - 1)
    xs = np.arange(-10.0, 10.0 + dx / 2, dx)
    m = len(xs)
    A = D.UpwindDifference(1, 3, dx, m, -1.0, offside=offside)
    u = torch.from_numpy(phi(xs, 0.0)).to("cuda")
    du = torch.empty_like(u)
    oA = po.Op(1, 1, 3, dx, m, offside=offside)
    oA.coeff = A.coefficients.copy()
    oA.coeff_kind = A.coeff_kind
    made = 0
    dt = 0.2 * dx
    for k in range(100):
        t = k * dt
        al = [0.0, 1.0, -6 * phi(-10.0, t), 0.0, -1.0]
        ar = [0.0, 1.0, -6 * phi(10.0, t), 0.0, -1.0]
        bc = D.GeneralBC(al, ar, dx, 3)
        c0 = plan_count()
        D.mul_(du, A, bc * u)
        made += plan_count() - c0
        fresh = D.plans_for(A * D.GeneralBC(al, ar, dx, 3), u)[0]
        ref_fresh = torch.empty_like(u)
        fresh.apply(ref_fresh, u)
        assert torch.equal(du, ref_fresh), k
        ref = po.ghost_apply([oA], [po.general(al, ar, dx, 3)], u.cpu().numpy())
        assert rel(du.cpu().numpy(), ref) <= 1e-12, k
        u = u - dt * du
    assert made <= 1   # one plan for the structure (UpwindDifference, GeneralBC of 5 taps)


def test_update_rejects_another_structure_and_keeps_values():
    n = 64
    h = 1.0 / (n + 1)
    L = D.CenteredDifference(2, 4, h, n)
    u = torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, n)).to("cuda")
    bc1 = D.RobinBC((1.0, 0.5, 0.2), (1.0, 0.5, -0.1), h, 1)
    p = D.plans_for(L * bc1, u)[0]
    ref = torch.empty_like(u)
    p.apply(ref, u)
    # same structure, new values: equal to a fresh plan
    bc2 = D.RobinBC((1.0, 0.7, -0.4), (0.9, 0.5, 0.3), h, 1)
    d2, keep2 = make_desc([L], (n,), _abi.F64, input_kind=_abi.INPUT_GHOST,
                          semantics=_abi.SEM_BPV, bcs={0: bc2})
    p.update(d2, keep2)
    got = torch.empty_like(u)
    p.apply(got, u)
    fresh = torch.empty_like(u)
    D.plans_for(L * bc2, u)[0].apply(fresh, u)
    assert torch.equal(got, fresh)
    # another structure (BC order 2: two taps a side): refused, values kept
    bc3 = D.RobinBC((1.0, 0.5, 0.2), (1.0, 0.5, -0.1), h, 2)
    d3, keep3 = make_desc([L], (n,), _abi.F64, input_kind=_abi.INPUT_GHOST,
                          semantics=_abi.SEM_BPV, bcs={0: bc3})
    with pytest.raises(_abi.DopError) as e:
        p.update(d3, keep3)
    assert e.value.code == _abi.EINVAL
    again = torch.empty_like(u)
    p.apply(again, u)
    assert torch.equal(again, fresh)


def test_update_coefficients_in_place():
    # new operator coefficients (update_coefficients!, DiffEqOperators.jl:67-75) of the same
    # structure: tables replaced, results equal to a fresh plan's
    n = 300
    h = 1.0 / (n + 1)
    rng = np.random.default_rng(4)
    u = torch.from_numpy(rng.uniform(-1, 1, (n, n))).to("cuda")
    Q = D.compose(*D.Dirichlet0BC(float, (n, n)))
    for k in range(5):
        v = rng.standard_normal(n)
        A = (D.CenteredDifference[1](2, 4, h, n) + D.UpwindDifference[2](1, 3, h, n, -v)) * Q
        got = A * u
        ref = torch.empty_like(u)
        D.plans_for((D.CenteredDifference[1](2, 4, h, n) + D.UpwindDifference[2](1, 3, h, n, -v)) * Q,
                    u)[0].apply(ref, u)
        assert torch.equal(got, ref), k
