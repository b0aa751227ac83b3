"""GPU parity at the full size of the 3-D non-uniform workload: the autotuned 512^3 plan of
per axis CenteredDifference(2,6,dx_k) + UpwindDifference(1,3,dx_k,-v_k) with Dirichlet0BC on
every face, applied in FP64 (<= 1e-12 normwise against the FP64 oracle) and in FP32 storage
(<= 1e-5 against the FP64 oracle on the same FP32-quantized input), through the C ABI."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu

N = 512


def workload(dtype):
    h = 1.0 / (N + 1)
    ops, oops = [], []
    for k in (1, 2, 3):
        dx = h * (1 + 0.5 * np.random.default_rng(10 + k).random(N + 1))
        v = np.random.default_rng(20 + k).standard_normal(N)
        ops += [D.CenteredDifference[k](2, 6, dx, N), D.UpwindDifference[k](1, 3, dx, N, -v)]
        oops += [po.Op(0, 2, 6, dx, N, axis=k - 1), po.Op(1, 1, 3, dx, N, axis=k - 1, coeff=-v)]
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    A = A * D.compose(*D.Dirichlet0BC(float, (N, N, N)))
    g = torch.Generator(device="cuda").manual_seed(3)
    u = (torch.rand((N, N, N), dtype=torch.float64, device="cuda", generator=g) * 2 - 1).to(dtype)
    return A, u, oops


def threads():
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail)


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-12)])
def test_nonuniform3d_512_autotuned_vs_oracle(dtype, tol):
    print(f"case: 512^3 non-uniform {dtype} autotuned vs oracle", flush=True)
    A, u, oops = workload(dtype)
    plan = D.plans_for(A, u)[0]
    variant = plan.autotune()
    du = torch.empty_like(u)
    plan.apply(du, u)
    got = du.double().cpu().numpy()
    uh = u.double().cpu().numpy()
    del du
    torch.cuda.empty_cache()
    po.set_threads(threads())
    try:
        ref = po.ghost_apply(oops, [po.dirichlet0()] * 3, uh)
    finally:
        po.set_threads(1)
    err = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"  variant {variant}: normwise rel err {err:.3e}", flush=True)
    assert err <= tol


def test_batched_1024_problem_mix_equals_single_applies():
    # the 1024-problem mix of the batched workload (1-D / 2-D, N log-uniform up to 65536 / 2048,
    # d in 1..4, a in {2,4,6,8}, Dirichlet / Neumann / Robin), FP32: the batched launch
    # equals every problem's own apply bit for bit
    from bench import batch_problem
    probs = [batch_problem(i, torch.float32) for i in range(1024)]
    print("case: 1024-problem batch vs single applies", flush=True)
    dus = [torch.empty_like(p[1]) for p in probs]
    b = D.Batch([p[0] for p in probs], [p[1] for p in probs], dus)
    b.apply()
    for i, p in enumerate(probs):
        ref = torch.empty_like(p[1])
        D.plans_for(p[0], p[1])[0].apply(ref, p[1])
        assert torch.equal(dus[i], ref), i


def test_batched_rk4_sample_vs_oracle():
    # 16 problems of the mix advanced by the graph-captured RK4 (3 steps) against the FP64
    # oracle's RK4 on the same FP32 initial values
    from bench import batch_problem
    idx = list(range(0, 1024, 64))
    probs = [batch_problem(i, torch.float32) for i in idx]
    u0 = [p[1].double().cpu().numpy() for p in probs]
    rk = D.RK4([p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs],
               [p[3] for p in probs])
    rk.step(3)
    torch.cuda.synchronize()
    for k, (p, u) in enumerate(zip(probs, u0)):
        A, sign, dt = p[0], p[2], p[3]
        terms = A.ops if isinstance(A, D.DiffEqOperatorCombination) else [A]
        oops, obcs = [], {}
        for G in terms:
            L = G.L
            o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
                      axis=L.N - 1, offside=L.offside)
            o.coeff = L.coefficients.copy()
            o.coeff_kind = L.coeff_kind
            oops.append(o)
        nd = u.ndim
        from diffeqoperators_amd.apply import _bcs_of
        bcs, _ = _bcs_of(terms[0].Q, nd, tuple(reversed(u.shape)))
        ob = [po.BC(bcs[a].a_l, bcs[a].b_l, bcs[a].a_r, bcs[a].b_r) for a in range(nd)]
        f = lambda v: sign * po.ghost_apply(oops, ob, v)
        for _ in range(3):
            k1 = f(u)
            k2 = f(u + 0.5 * dt * k1)
            k3 = f(u + 0.5 * dt * k2)
            k4 = f(u + dt * k3)
            u = u + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        got = p[1].double().cpu().numpy()
        err = np.abs(got - u).max() / np.abs(u).max()
        assert err < 1e-5, (idx[k], err)


def test_served_200_request_mix_sample_equals_alone():
    # the served workload's 200 requests (Poisson arrivals at 2000/s, 64^2..2048^2, 100 RK4
    # steps): a sample of 10 results equals the same request served alone, bit for bit
    from bench import served_requests
    reqs = served_requests(200, 2000.0)
    plans = [D.plan_for_shape(A, (n, n), torch.float32) for A, n, _, _, _ in reqs]
    srv = D.Server(8 << 30, max_active=4096, chunk=16)
    for (A, n, u0, dt, rel), plan in zip(reqs, plans):
        srv.submit(plan, u0, 1.0, dt, 100, rel)
    print("case: 200-request served mix", flush=True)
    srv.run()
    for i in range(0, 200, 20):
        A, n, u0, dt, _ = reqs[i]
        alone = D.Server(1 << 30, max_active=4, chunk=16)
        j = alone.submit(plans[i], u0, 1.0, dt, 100, 0.0)
        alone.run()
        assert torch.equal(srv.results[i], alone.results[j]), i
