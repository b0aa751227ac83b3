"""GPU: the tiled interior kernels (register ring + LDS) against the generic per-point
kernel, which they must match bit for bit (same FMA sequence per point), over shapes that
exercise partial x segments, unaligned rows, narrow boxes, every window width and the
accumulate path; plus the XCD / chunk tile-shape candidates of the autotuner."""
import os

import numpy as np
import pytest
import torch

import diffeqoperators_amd as D

pytestmark = pytest.mark.gpu


def build(shape, K_half, dtype, seed, uniform=True, upwind=False):
    rng = np.random.default_rng(seed)
    nd = len(shape)
    jsize = tuple(reversed(shape))
    ops = []
    for N in range(1, nd + 1):
        n = jsize[N - 1]
        h = 1.0 / (n + 1)
        dx = h if uniform in (True, "mixed") else h * (1 + 0.5 * rng.random(n + 1))
        # CenteredDifference(2, 2*K_half) has a 2*K_half+1 point interior window
        ops.append(D.CenteredDifference[N](2, 2 * K_half, dx, n))
        if upwind or (uniform == "mixed" and N == 1):
            # "mixed": per-row weights along x only (vector coefficients), uniform elsewhere
            ops.append(D.UpwindDifference[N](1, 3, dx, n, rng.standard_normal(n)))
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.2), tuple(1.0 / (s + 1) for s in jsize), 1, jsize))
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    return A * Q, u


def generic_of(A, u):
    os.environ["DOP_FORCE_GENERIC"] = "1"
    try:
        A2 = D.DiffEqOperatorCombination(A.ops) if isinstance(A, D.DiffEqOperatorCombination) else A
        A2._plans = {}
        return D.plans_for(A2, u)[0]
    finally:
        os.environ.pop("DOP_FORCE_GENERIC", None)


CASES_2D = [((64, 64), 1), ((257, 130), 2), ((100, 333), 3), ((4096, 64), 2), ((37, 1000), 4),
            ((200, 190), 5), ((301, 77), 6), ((20, 2048), 2)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape,kh", CASES_2D)
@pytest.mark.parametrize("uniform", [True, False, "mixed"])
def test_tiled_2d_bitwise(shape, kh, dtype, uniform):
    A, u = build(shape, kh, dtype, seed=kh, uniform=uniform, upwind=not uniform and kh == 3)
    print(f"case: tiled2d {shape} K={2*kh+1} {dtype}", flush=True)
    tiled = D.plans_for(A, u)[0]
    gen = generic_of(A, u)
    a = torch.empty_like(u)
    b = torch.empty_like(u)
    tiled.apply(a, u)
    gen.apply(b, u)
    assert torch.equal(a, b)
    # accumulate
    a2 = a.clone()
    b2 = b.clone()
    tiled.apply(a2, u, accumulate=True)
    gen.apply(b2, u, accumulate=True)
    assert torch.equal(a2, b2)


CASES_3D = [((32, 32, 32), 1), ((40, 48, 64), 2), ((33, 29, 70), 3), ((64, 64, 256), 3),
            ((17, 100, 36), 4), ((96, 96, 96), 3), ((12, 40, 600), 3), ((19, 23, 532), 2)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape,kh", CASES_3D)
@pytest.mark.parametrize("uniform", [True, False, "mixed"])
def test_tiled_3d_bitwise(shape, kh, dtype, uniform):
    A, u = build(shape, kh, dtype, seed=10 + kh, uniform=uniform, upwind=not uniform and kh == 3)
    print(f"case: tiled3d {shape} K={2*kh+1} {dtype} uniform={uniform}", flush=True)
    tiled = D.plans_for(A, u)[0]
    gen = generic_of(A, u)
    a = torch.empty_like(u)
    b = torch.empty_like(u)
    tiled.apply(a, u)
    gen.apply(b, u)
    assert torch.equal(a, b)
    # accumulate (the x boundary rows the tiles evaluate take the same store path)
    a2 = a.clone()
    b2 = b.clone()
    tiled.apply(a2, u, accumulate=True)
    gen.apply(b2, u, accumulate=True)
    assert torch.equal(a2, b2)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape", [(257, 130), (40, 48, 64), (33, 29, 70), (64, 64, 256)])
def test_tiled_wide_accumulation_bitwise(shape, dtype):
    # third derivatives: DOP_ACC_AUTO sums f32 values in f64 (DOP_ACC_F32_WIDE); the tiles
    # and the 3-D shell run that arithmetic and must match the generic kernel bit for bit
    rng = np.random.default_rng(21)
    jsize = tuple(reversed(shape))
    ops = []
    for N in range(1, len(shape) + 1):
        n = jsize[N - 1]
        dx = (1.0 / (n + 1)) * (1 + 0.5 * rng.random(n + 1))
        ops += [D.CenteredDifference[N](3, 4, dx, n), D.CenteredDifference[N](2, 2, dx, n)]
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.1), (1.0, 0.5, -0.2), tuple(1.0 / (s + 1) for s in jsize), 1, jsize))
    A = A * Q
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    print(f"case: wide {shape} {dtype}", flush=True)
    tiled = D.plans_for(A, u)[0]
    gen = generic_of(A, u)
    a = torch.empty_like(u)
    b = torch.empty_like(u)
    tiled.apply(a, u)
    gen.apply(b, u)
    assert torch.equal(a, b)
    a2, b2 = a.clone(), b.clone()
    tiled.apply(a2, u, accumulate=True)
    gen.apply(b2, u, accumulate=True)
    assert torch.equal(a2, b2)


def test_autotune_candidates_agree():
    A, u = build((512, 512), 2, torch.float64, seed=3)
    p = D.plans_for(A, u)[0]
    ref = torch.empty_like(u)
    p.apply(ref, u)
    choice = p.autotune()
    assert choice >= 0
    out = torch.empty_like(u)
    p.apply(out, u)
    assert torch.equal(out, ref)


def plan_with_variant(A, u, v):
    os.environ["DOP_TILE_VARIANT"] = str(v)
    try:
        A2 = D.DiffEqOperatorCombination(A.ops)
        return D.plans_for(A2, u)[0]
    finally:
        os.environ.pop("DOP_TILE_VARIANT", None)


def n_variants(shape, dtype, kh):
    """Tile-shape candidates of the tiled path (dop_fast.cpp: fast_variant_count); kh is the
    window's half-width."""
    n0 = shape[-1]
    vx = 2 if dtype == torch.float64 else 4
    vx = vx if n0 % vx == 0 else 1
    if len(shape) == 2:
        # chunk x prefetch x remap (16) x points per lane (2, when rows allow) x marches (2)
        # x segments (overlapping; edge-loaded when kh <= vx)
        base = 64 if vx > 1 and n0 % (2 * vx) == 0 else 32
        return base * (2 if kh <= vx else 1)
    # overlapping segments (40 / 20) and edge-loaded segments (32 / 16)
    return 72 if vx > 1 else 36


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("shape,kh", [((300, 260), 2), ((40, 1040), 2), ((50, 301), 1),
                                      ((40, 72, 100), 3), ((33, 96), 5),
                                      ((21, 45, 532), 3), ((14, 30, 266), 4)])
def test_every_tile_variant_bitwise(shape, kh, dtype):
    nvar = n_variants(shape, dtype, kh)
    A, u = build(shape, kh, dtype, seed=5, uniform=False, upwind=(kh == 3))
    ref = torch.empty_like(u)
    generic_of(A, u).apply(ref, u)
    for v in range(nvar):
        print(f"case: variant {v} {shape} {dtype}", flush=True)
        out = torch.empty_like(u)
        plan_with_variant(A, u, v).apply(out, u)
        assert torch.equal(out, ref), v
