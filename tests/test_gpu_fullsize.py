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
