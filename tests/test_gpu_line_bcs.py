"""GPU parity with per-line boundary conditions, MultiDimBC{dim}(BC::Array)
(multi_dim_bc_operators.jl:93-95): the fused apply (tiled interior + generic shell) and the
device concretization against the FP64 oracle with the same per-line BCs; FP64 1e-12, FP32
storage 1e-5 (normwise)."""
import numpy as np
import pytest
import torch

import pyoracle as po
import diffeqoperators_amd as D
from test_host_line_bcs import bc_array, obc, oop

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300)


def _problem(rng, shape, periodic=0.0):
    jsize = tuple(reversed(shape))
    nd = len(shape)
    ops, oops, Qs, obcs = [], [], [], []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        dx = 0.05 * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](2, 4, dx, n), D.UpwindDifference[N](1, 3, dx, n, -rng.standard_normal(n))]
        oops += [oop(ops[-2]), oop(ops[-1])]
        other = tuple(jsize[k] for k in range(nd) if k != N - 1)
        arr = bc_array(rng, other, 0.05, periodic)
        Qs.append(D.MultiDimBC[N](arr))
        obcs.append(po.BC.per_line([obc(b) for b in arr.flatten(order="F")]))
    return sum(ops[1:], ops[0]) * D.compose(*Qs), oops, obcs


@pytest.mark.parametrize("shape", [(96, 130), (24, 20, 70)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("periodic", [0.0, 0.5])
def test_per_line_apply_matches_oracle(rng, shape, dtype, periodic):
    # periodic = 0.5: the multi_dim_bc_test.jl pattern, Robin and PeriodicBC lines mixed
    G, oops, obcs = _problem(rng, shape, periodic)
    print(f"[line-bc] {shape} {dtype}", flush=True)
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda", dtype)
    got = (G * u).double().cpu().numpy()
    ref = po.ghost_apply(oops, obcs, u.double().cpu().numpy())
    assert rel(got, ref) <= (1e-12 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("periodic", [0.0, 0.5])
def test_per_line_concretization_matches_apply(rng, periodic):
    import scipy.sparse as sp
    shape = (14, 19)
    G, oops, obcs = _problem(rng, shape, periodic)
    Al, Ab = D.sparse(G, tuple(reversed(shape)))
    S = sp.csr_matrix((Al.values().cpu().numpy(), Al.col_indices().cpu().numpy(),
                       Al.crow_indices().cpu().numpy()), shape=tuple(Al.shape))
    u = rng.uniform(-1, 1, shape)
    ref = po.ghost_apply(oops, obcs, u).ravel()
    got = S @ u.ravel() + Ab.cpu().numpy()
    assert rel(got, ref) <= 1e-12
