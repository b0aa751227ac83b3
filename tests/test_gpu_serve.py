"""GPU: the continuous-batching service.  Requests of different grid sizes released over
time must give, in the mix, exactly the result each gives alone (same kernels, same
arithmetic: bitwise), match an FP64 oracle RK4 within 1e-4 (FP32 storage), respect the
slot and HBM-budget limits, and be admitted in FIFO order."""
import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu


def advdiff(n, seed):
    """(C{1}(2,4) + C{2}(2,4) + U{1}(1,3,-vx) + U{2}(1,3,-vy)) * Dirichlet0, n x n."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n + 1)
    vx, vy = rng.standard_normal(n), rng.standard_normal(n)
    ops = [D.CenteredDifference[1](2, 4, h, n), D.CenteredDifference[2](2, 4, h, n),
           D.UpwindDifference[1](1, 3, h, n, -vx), D.UpwindDifference[2](1, 3, h, n, -vy)]
    A = ops[0] + ops[1] + ops[2] + ops[3]
    Q = D.compose(*D.Dirichlet0BC(float, (n, n)))
    u0 = torch.from_numpy(rng.uniform(-1, 1, (n, n)).astype(np.float32)).pin_memory()
    wsum = sum(float(np.abs(L.stencil_coefs).sum()) for L in ops) * (1 + np.abs(np.r_[vx, vy]).max())
    return A * Q, ops, u0, 0.4 / wsum


def run_server(reqs, budget=1 << 28, max_active=3, nsteps=8):
    srv = D.Server(budget, max_active=max_active, chunk=16)
    ids = []
    for A, _, u0, dt, rel, n in reqs:
        plan = D.plan_for_shape(A, (n, n), torch.float32)
        ids.append(srv.submit(plan, u0, 1.0, dt, nsteps, rel))
    srv.run()
    return srv, [srv.results[i].clone() for i in ids]


@pytest.fixture(scope="module")
def requests():
    sizes = [64, 200, 96, 131, 250, 77, 180, 64, 300, 111]
    out = []
    for i, n in enumerate(sizes):
        A, ops, u0, dt = advdiff(n, 100 + i)
        out.append((A, ops, u0, dt, 0.0005 * i, n))
    return out


def test_mix_equals_alone(requests):
    srv, mixed = run_server(requests)
    st = srv.stats()
    assert st["peak_active"] <= 3
    t = srv.times()
    assert np.all(t[:, 1] >= t[:, 0] - 1e-9) and np.all(t[:, 2] >= t[:, 1])
    # FIFO admission: admission times are non-decreasing in release order
    order = np.argsort(t[:, 0], kind="stable")
    assert np.all(np.diff(t[order, 1]) >= -1e-9)
    for k, r in enumerate(requests):
        _, alone = run_server([r[:4] + (0.0, r[5])])
        assert torch.equal(mixed[k], alone[0]), k


def test_budget_limits_residency(requests):
    n_max = max(r[5] for r in requests)
    per = 4 * (n_max * n_max * 4 + 256)
    srv, _ = run_server(requests, budget=2 * per + 4096, max_active=64)
    st = srv.stats()
    assert st["peak_bytes"] <= 2 * per + 4096
    assert st["peak_active"] >= 2


def test_served_matches_oracle_rk4(requests):
    A, ops, u0, dt, _, n = requests[0]
    _, res = run_server([(A, ops, u0, dt, 0.0, n)], nsteps=6)
    bc = po.dirichlet0()
    oops = []
    for L in ops:
        o = po.Op(int(L.wind), L.derivative_order, L.approximation_order, L.dx, L.len,
                  axis=L.N - 1, offside=L.offside)
        o.coeff = L.coefficients.copy()
        o.coeff_kind = L.coeff_kind
        oops.append(o)
    f = lambda v: po.ghost_apply(oops, [bc, bc], v)
    u = u0.numpy().astype(np.float64)
    for _ in range(6):
        k1 = f(u)
        k2 = f(u + 0.5 * dt * k1)
        k3 = f(u + 0.5 * dt * k2)
        k4 = f(u + dt * k3)
        u = u + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    got = res[0].numpy().astype(np.float64)
    err = np.abs(got - u).max() / np.abs(u).max()
    print(f"served vs FP64 oracle RK4 (6 steps, FP32 storage): {err:.3e}")
    assert err < 1e-4


def test_live_submission_while_running(requests):
    # requests submitted from another thread while run() serves: each completes, after its
    # release, with the result it gives alone
    import threading
    srv = D.Server(1 << 28, max_active=4, chunk=16)
    srv.open()
    th = threading.Thread(target=srv.run)
    th.start()
    ids, plans = [], []
    import time
    try:
        for k, (A, _, u0, dt, _, n) in enumerate(requests[:6]):
            plan = D.plan_for_shape(A, (n, n), torch.float32)
            plans.append(plan)
            ids.append(srv.submit(plan, u0, 1.0, dt, 8))   # released now
            time.sleep(0.002)
    finally:
        srv.close()
        th.join(timeout=60)
    assert not th.is_alive()
    t = srv.times()
    assert np.all(t[:, 2] >= t[:, 0])
    assert t[-1, 0] > 0.0     # the last ones were submitted after the run had started
    for k, i in enumerate(ids):
        r = requests[k]
        _, alone = run_server([r[:4] + (0.0, r[5])])
        assert torch.equal(srv.results[i], alone[0]), k


def test_small_requests_keep_their_own_pace():
    # two large grids and eight small ones released together: the small ones run in their own
    # size class and finish before either large one
    reqs = [advdiff(1100, 900), advdiff(1100, 901)] + [advdiff(64, 910 + k) for k in range(8)]
    srv = D.Server(1 << 30, max_active=64, chunk=16)
    ids = []
    for (A, _, u0, dt), n in zip(reqs, [1100, 1100] + [64] * 8):
        ids.append(srv.submit(D.plan_for_shape(A, (n, n), torch.float32), u0, 1.0, dt, 60, 0.0))
    srv.run()
    t = srv.times()
    assert t[2:, 2].max() < t[:2, 2].min()


def test_budget_is_checked_on_the_rounded_state():
    A, _, u0, dt = advdiff(50, 77)
    plan = D.plan_for_shape(A, (50, 50), torch.float32)
    state = 4 * ((50 * 50 * 4 + 255) // 256 * 256)
    small = D.Server(state - 1, max_active=4)       # rounds down below the state: refused
    with pytest.raises(D._abi.DopError) as e:
        small.submit(plan, u0, 1.0, dt, 2, 0.0)
    assert e.value.code == D._abi.EBUDGET
    ok = D.Server(state + 100, max_active=4)          # rounds down to exactly the state
    ok.submit(plan, u0, 1.0, dt, 2, 0.0)
    ok.submit(plan, u0, 1.0, dt, 2, 0.0)              # waits for the first one's memory
    ok.run()
    assert ok.stats()["peak_active"] == 1
