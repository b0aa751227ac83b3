#!/usr/bin/env python3
"""Throughput of the fused operator application on one MI355X.

Default workload (`--config nonuniform3d`): the 3-D non-uniform 512^3 field, per axis k
CenteredDifference{k}(2,6,dx_k,512) + UpwindDifference{k}(1,3,dx_k,512,-v_k) summed and composed
with Dirichlet0BC on every face, FP32 storage and accumulation: one step = one (sum L)*Q*u.
The FP64 and BF16 storage variants are timed beside it, each with its error against the FP64
oracle on the same quantized input, and the oracle is timed on the same workload (the CPU
baseline).  Prints ONE JSON line.

Other workloads are reported in the same line under "also" (`--no-also` skips them), or
selected with `--config laplacian2d|batched|served`.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "diffeqoperators.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "stencil Gpts/s + % of HBM peak; solves/s + p99 latency, served vs plain"
HBM_PEAK_GBS = 8000.0


def _sync():
    import torch
    torch.cuda.synchronize()


def time_applies(fn, steps, warmup):
    """Wall time and HIP-event time (on torch's current stream, where libdop launches).  The
    garbage collector is off inside the timed loop (as in timeit): a full collection stalls the
    host for tens of milliseconds, long enough to drain the launch queue and leave the GPU idle
    inside the event span."""
    import torch
    for _ in range(warmup):
        fn()
    _sync()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    was = gc.isenabled()
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        e0.record(s)
        for _ in range(steps):
            fn()
        e1.record(s)
        _sync()
        wall = time.perf_counter() - t0
    finally:
        if was:
            gc.enable()
    return wall, e0.elapsed_time(e1) / 1e3


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/<name>_pmc.txt)."""
    path = os.path.join(ROOT, "profiles", f"{name}_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


# ------------------------------------------------------------------------------------------
def laplacian2d(n=4096):
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    h = 1.0 / (n + 1)
    Lx = D.CenteredDifference[1](2, 4, h, n)
    Ly = D.CenteredDifference[2](2, 4, h, n)
    Q = D.compose(*D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), (h, h), 1, (n, n)))
    A = (Lx + Ly) * Q
    rng = np.random.default_rng(1)
    u = torch.from_numpy(rng.uniform(-1.0, 1.0, (n, n))).to("cuda")
    return A, u, [Lx, Ly], Q


def bench_laplacian2d(steps, warmup):
    import torch
    import diffeqoperators_amd as D
    n = 4096
    A, u, _, _ = laplacian2d(n)
    plan = D.plans_for(A, u)[0]
    # a tile variant pinned by DOP_TILE_VARIANT (profiling runs) is kept
    variant = plan.variant[0] if os.environ.get("DOP_TILE_VARIANT") else plan.autotune()
    du = torch.empty_like(u)
    wall, ev = time_applies(lambda: plan.apply(du, u), steps, warmup)
    pts = n * n
    alg_bytes = 2 * 8 * pts
    kern_s = ev / steps
    achieved = alg_bytes / kern_s / 1e9
    return dict(
        value=pts * steps / wall / 1e9, ms_per_step=wall / steps * 1e3, kernel_ms=kern_s * 1e3,
        roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 4),
                      traffic=read_traffic("apply_laplacian2d_4096_f64")),
        alg_bytes=alg_bytes, variant=variant)


def bench_laplacian2d_f32(steps, warmup):
    """The same 2-D workload with FP32 storage and FP32 accumulation, and its normwise error
    against the FP64 apply of the same (rounded) field."""
    import torch
    import diffeqoperators_amd as D
    n = 4096
    A, u64, _, _ = laplacian2d(n)
    u = u64.float()
    plan = D.plans_for(A, u)[0]
    variant = plan.autotune()
    du = torch.empty_like(u)
    wall, ev = time_applies(lambda: plan.apply(du, u), steps, warmup)
    ref = torch.empty_like(u64)
    D.plans_for(A, u64)[0].apply(ref, u.double())
    err = ((du.double() - ref).abs().max() / ref.abs().max()).item()
    pts = n * n
    kern_s = ev / steps
    achieved = 2 * 4 * pts / kern_s / 1e9
    return dict(gpts_s=round(pts * steps / wall / 1e9, 2), kernel_ms=round(kern_s * 1e3, 4),
                achieved_GBs=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4),
                max_rel_err_vs_f64=float(f"{err:.3e}"), tile_variant=variant)


def bench_widened(steps=50, warmup=5):
    """Throughput of the callers beside the application path: Gradient / Divergence / Curl of
    a 256^3 FP32 field (padded input), nonlinear_diffusion! over 2^24 FP64 knots, and the 2-D
    workload with one Robin BC per line along x.  Algorithmic bytes: every input array read once
    and every output written once."""
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    out = {}

    def rec(name, fn, pts, alg_bytes):
        wall, ev = time_applies(fn, steps, warmup)
        t = ev / steps
        ach = alg_bytes / t / 1e9
        out[name] = dict(gpts_s=round(pts / t / 1e9, 2), kernel_ms=round(t * 1e3, 4),
                         achieved_GBs=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4))

    n = 256
    h = 1.0 / (n + 1)
    g = torch.Generator(device="cuda").manual_seed(5)
    M = torch.rand((n + 2,) * 3, device="cuda", generator=g) * 2 - 1
    V = torch.rand((3,) + (n + 2,) * 3, device="cuda", generator=g) * 2 - 1
    G = D.Gradient(4, (h, h, h), (n, n, n))
    Dv = D.Divergence(4, (h, h, h), (n, n, n))
    Cu = D.Curl(4, (h, h, h), (n, n, n))
    rec("gradient_256_f32", lambda: G * M, n ** 3, 4 * ((n + 2) ** 3 + 3 * n ** 3))
    rec("divergence_256_f32", lambda: Dv * V, n ** 3, 4 * (3 * (n + 2) ** 3 + n ** 3))
    rec("curl_256_f32", lambda: Cu * V, n ** 3, 4 * (3 * (n + 2) ** 3 + 3 * n ** 3))
    del M, V
    torch.cuda.empty_cache()

    m = 1 << 24
    x = torch.linspace(-5.0, 5.0, m + 2, dtype=torch.float64, device="cuda")
    q = 1.0 / torch.sqrt(x * x + 1.0)
    pk = q ** -2
    du = torch.empty(m, dtype=torch.float64, device="cuda")
    dxn = 10.0 / (m + 1)
    rec("nonlinear_diffusion_2e24_f64",
        lambda: D.nonlinear_diffusion_(du, 1, 1, 2, pk, q, dxn, m), m, 8 * (2 * (m + 2) + m))
    del x, q, pk, du
    torch.cuda.empty_cache()

    nn = 4096
    hh = 1.0 / (nn + 1)
    rng = np.random.default_rng(7)
    lines = np.empty((nn,), dtype=object)
    for i in range(nn):
        lines[i] = D.RobinBC((1.0, float(rng.uniform(0.2, 1.0)), float(rng.standard_normal())),
                             (1.0, float(rng.uniform(0.2, 1.0)), float(rng.standard_normal())),
                             hh, 1)
    Q = D.compose(D.MultiDimBC[1](lines),
                  D.MultiDimBC[2](D.RobinBC((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), hh, 1), (nn, nn)))
    A = (D.CenteredDifference[1](2, 4, hh, nn) + D.CenteredDifference[2](2, 4, hh, nn)) * Q
    u = torch.from_numpy(rng.uniform(-1.0, 1.0, (nn, nn))).to("cuda")
    plan = D.plans_for(A, u)[0]
    plan.autotune()
    dv = torch.empty_like(u)
    rec("laplacian2d_4096_f64_per_line_robin", lambda: plan.apply(dv, u), nn * nn, 2 * 8 * nn * nn)
    return out


def bench_copy(steps=50, warmup=5, sizes_mib=(64, 128, 256, 512, 1024)):
    """A plain device copy (torch copy_) of each size to a second buffer of that size: the rate
    the memory system gives a pure stream.  Working sets that fit MI355X's 256 MB Infinity Cache
    (a copy of up to 128 MiB) are served from it when repeated, as the 2-D 4096^2 applies are;
    larger ones run at the HBM rate that bounds the 512^3 apply."""
    import torch
    out = {}
    for mib in sizes_mib:
        a = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        _, ev = time_applies(lambda: b.copy_(a), steps, warmup)
        t = ev / steps
        gbs = 2 * (mib << 20) / t / 1e9
        out[f"{mib}MiB"] = dict(us=round(t * 1e6, 1), read_write_GBs=round(gbs, 1),
                                frac=round(gbs / HBM_PEAK_GBS, 4))
        del a, b
        torch.cuda.empty_cache()
    return out


def cpu_baseline_laplacian2d(seconds):
    """The oracle (FP64 C restatement of the reference's per-line loops) on the same
    workload, one thread, on a bounded sample: whole applies until `seconds` elapse."""
    import numpy as np
    import pyoracle as po
    n = 4096
    h = 1.0 / (n + 1)
    ops = [po.Op(0, 2, 4, h, n, axis=0), po.Op(0, 2, 4, h, n, axis=1)]
    bc = po.robin((1.0, 0.5, 0.0), (1.0, 0.5, 0.0), h, 1)
    u = np.random.default_rng(1).uniform(-1.0, 1.0, (n, n))
    def run(threads, budget):
        po.set_threads(threads)
        reps, t0 = 0, time.perf_counter()
        while True:
            po.ghost_apply(ops, [bc, bc], u)
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget:
                return reps, el

    reps, el = run(1, seconds)
    # the same loops over the host's cores (OpenMP over lines), at most 16 threads
    cores = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                       os.cpu_count() or 1))
    reps_m, el_m = run(cores, max(2.0, seconds / 2))
    po.set_threads(1)
    return dict(value=round(n * n * reps / el / 1e9, 4), unit="Gpts/s", cores=1, kind="port",
                sample=f"{reps} full 4096x4096 FP64 applies of the oracle (per-line loops, one "
                       f"temporary per operator), {el:.1f} s",
                all_cores=dict(value=round(n * n * reps_m / el_m / 1e9, 4), cores=cores,
                               sample=f"{reps_m} applies, {el_m:.1f} s, OpenMP over lines"))


# ------------------------------------------------------------------------------------------
def nonuniform3d(n, dtype):
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    h = 1.0 / (n + 1)
    ops = []
    for k in (1, 2, 3):
        dx = h * (1 + 0.5 * np.random.default_rng(10 + k).random(n + 1))
        v = np.random.default_rng(20 + k).standard_normal(n)
        ops += [D.CenteredDifference[k](2, 6, dx, n), D.UpwindDifference[k](1, 3, dx, n, -v)]
    Q = D.compose(*D.Dirichlet0BC(float, (n, n, n)))
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    A = A * Q
    g = torch.Generator(device="cuda").manual_seed(3)
    u = (torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1).to(dtype)
    return A, u


def bench_slab(steps=50, warmup=5, nz=64, halo=3):
    """One rank's share of the 512^3 FP32 workload sharded in z over 8 GPUs, at its real shape:
    512 x 512 x (64 + 2*3) planes (the 3 halo planes per side are part of the array).  Applied
    as a 70-plane problem, whose 3 edge planes per side are evaluated as boundary rows where a
    rank would read halo planes: an upper bound on the per-rank compute."""
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    n = 512
    h = 1.0 / (n + 1)
    sizes = (n, n, nz + 2 * halo)
    ops = []
    for k in (1, 2, 3):
        m = sizes[k - 1]
        dx = h * (1 + 0.5 * np.random.default_rng(10 + k).random(m + 1))
        v = np.random.default_rng(20 + k).standard_normal(m)
        ops += [D.CenteredDifference[k](2, 6, dx, m), D.UpwindDifference[k](1, 3, dx, m, -v)]
    A = ops[0]
    for L in ops[1:]:
        A = A + L
    A = A * D.compose(*D.Dirichlet0BC(float, sizes))
    g = torch.Generator(device="cuda").manual_seed(3)
    u = torch.rand(tuple(reversed(sizes)), device="cuda", generator=g) * 2 - 1
    plan = D.plans_for(A, u)[0]
    variant = plan.autotune()
    du = torch.empty_like(u)
    wall, ev = time_applies(lambda: plan.apply(du, u), steps, warmup)
    t = ev / steps
    pts = n * n * nz
    ach = 2 * 4 * u.numel() / t / 1e9
    return dict(shape=list(reversed(sizes)), owned_points=pts, kernel_us=round(t * 1e6, 1),
                owned_gpts_s=round(pts / t / 1e9, 2), achieved_GBs=round(ach, 1),
                frac=round(ach / HBM_PEAK_GBS, 4), tile_variant=variant,
                halo_face_MB=round(n * n * halo * 4 / 1e6, 2))


def nonuniform3d_oracle_ops(n):
    """The oracle's operators of the 512^3 workload (the same seeded dx and wind)."""
    import numpy as np
    import pyoracle as po
    h = 1.0 / (n + 1)
    ops = []
    for k in (1, 2, 3):
        dx = h * (1 + 0.5 * np.random.default_rng(10 + k).random(n + 1))
        v = np.random.default_rng(20 + k).standard_normal(n)
        ops += [po.Op(0, 2, 6, dx, n, axis=k - 1), po.Op(1, 1, 3, dx, n, axis=k - 1, coeff=-v)]
    return ops, [po.dirichlet0()] * 3


def host_threads():
    """(threads used, host cores): the cores this process may run on, capped by
    OMP_NUM_THREADS when the environment sets it (the share of the host given to one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    use = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    return max(1, use), os.cpu_count() or avail


def rel_errors(du, ref):
    """normwise |du - ref|_inf / |ref|_inf and pointwise max |du - ref| / max(|ref|,
    1e-3 |ref|_inf)"""
    import numpy as np
    d = np.abs(du - ref)
    nr = float(np.abs(ref).max())
    return float(d.max() / nr), float((d / np.maximum(np.abs(ref), 1e-3 * nr)).max())


def bench_nonuniform3d(steps, warmup, n=512, errors=True, cpu_seconds=12.0,
                       types=("f32", "f64", "bf16")):
    """The 512^3 non-uniform workload in FP64 / FP32 / BF16 storage (FP32 accumulation for the
    narrow types), each with its error against the FP64 oracle on the same quantized input, and
    the oracle itself timed on the same workload (the CPU baseline)."""
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    out, dus, us = {}, {}, {}
    pts = n ** 3
    for name, dt, sb in (("f32", torch.float32, 4), ("f64", torch.float64, 8),
                         ("bf16", torch.bfloat16, 2)):
        if name not in types:
            continue
        A, u = nonuniform3d(n, dt)
        plan = D.plans_for(A, u)[0]
        # a tile variant pinned by DOP_TILE_VARIANT (profiling runs) is kept
        variant = plan.variant[0] if os.environ.get("DOP_TILE_VARIANT") else plan.autotune()
        du = torch.empty_like(u)
        wall, ev = time_applies(lambda: plan.apply(du, u), steps, warmup)
        kern = ev / steps
        ach = 2 * sb * pts / kern / 1e9
        out[name] = dict(gpts_s=round(pts * steps / wall / 1e9, 2),
                         ms_per_step=round(wall / steps * 1e3, 5),
                         kernel_ms=round(kern * 1e3, 4), achieved_GBs=round(ach, 1),
                         frac=round(ach / HBM_PEAK_GBS, 4), tile_variant=variant)
        if errors:
            dus[name] = du.double().cpu().numpy()
            us[name] = u.double().cpu().numpy()
        del A, u, du, plan
        torch.cuda.empty_cache()
    if not errors or "f64" not in us:
        return out, None
    import pyoracle as po
    ops, bcs = nonuniform3d_oracle_ops(n)
    # CPU baseline: the oracle (FP64 per-line loops, one temporary per operator) on the FP64
    # input, one thread, whole applies; then the same on every core this process may use
    reps, t0 = 0, time.perf_counter()
    po.set_threads(1)
    while True:
        ref64 = po.ghost_apply(ops, bcs, us["f64"])
        reps += 1
        el1 = time.perf_counter() - t0
        if el1 >= cpu_seconds:
            break
    thr, host_cores = host_threads()
    po.set_threads(thr)
    reps_m, t0 = 0, time.perf_counter()
    while True:
        po.ghost_apply(ops, bcs, us["f64"])
        reps_m += 1
        el_m = time.perf_counter() - t0
        if el_m >= max(2.0, cpu_seconds / 3):
            break
    for name in ("f64", "f32", "bf16"):
        if name not in us:
            continue
        ref = ref64 if name == "f64" else po.ghost_apply(ops, bcs, us[name])
        nw, pw = rel_errors(dus[name], ref)
        out[name]["max_rel_err_vs_f64"] = float(f"{nw:.3e}")
        out[name]["pointwise_rel_err_vs_f64"] = float(f"{pw:.3e}")
    po.set_threads(1)
    cpu = dict(value=round(pts * reps / el1 / 1e9, 5), unit="Gpts/s", cores=1, kind="port",
               sample=f"{reps} full 512^3 FP64 applies of the oracle (per-line loops of "
                      f"convolutions.jl, one temporary per operator), {el1:.1f} s, one thread",
               all_cores=dict(value=round(pts * reps_m / el_m / 1e9, 5), cores=thr,
                              host_cores=host_cores,
                              sample=f"{reps_m} applies, {el_m:.1f} s, OpenMP over lines"))
    return out, cpu


# ------------------------------------------------------------------------------------------
def batch_problem(i, dtype):
    """Problem i of the 1024-problem mix: 1-D or 2-D, N log-uniform, d in 1..4,
    a in {2,4,6,8}, Dirichlet / Neumann / Robin; RHS sign (-1)^(d/2+1) (even d), -1 (odd)."""
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    rng = np.random.default_rng(1000 + i)
    ndim = int(rng.integers(1, 3))
    d = int(rng.integers(1, 5))
    a = int(rng.choice([2, 4, 6, 8]))
    hi = 65536 if ndim == 1 else 2048
    n = int(np.exp(rng.uniform(np.log(64), np.log(hi))))
    h = 1.0 / (n + 1)
    g = float(rng.uniform(-1, 1))
    kind = int(rng.integers(0, 3))
    atom = (D.DirichletBC(g, -g) if kind == 0 else
            D.NeumannBC((g, 0.5 * g), h, 1) if kind == 1 else
            D.RobinBC((1.0, 0.5, g), (1.0, 0.5, -g), h, 1))
    if ndim == 1:
        ops = [D.CenteredDifference(d, a, h, n)]
        A = ops[0] * atom
        shape = (n,)
    else:
        ops = [D.CenteredDifference[1](d, a, h, n), D.CenteredDifference[2](d, a, h, n)]
        A = (ops[0] + ops[1]) * D.compose(*D.MultiDimBC(atom, (n, n)))
        shape = (n, n)
    sign = -1.0 if d % 2 == 1 else float((-1) ** (d // 2 + 1))
    wsum = sum(float(np.abs(L.stencil_coefs).sum()) for L in ops)
    u = torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda").to(dtype)
    return A, u, sign, 0.5 / wsum


def bench_batched(steps, nprob=1024, plain_sample=48, plain_steps=4):
    """1024 independent problems, graph-captured RK4: each stage is a few grouped launches (one
    per kind of problem), spread over four streams inside the graph."""
    import torch
    import diffeqoperators_amd as D
    probs = [batch_problem(i, torch.float32) for i in range(nprob)]
    pts = sum(p[1].numel() for p in probs)
    t0 = time.perf_counter()
    rk = D.RK4([p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs],
               [p[3] for p in probs])
    t_create = time.perf_counter() - t0
    rk.step(2)
    _sync()
    wall, ev = time_applies(lambda: rk.step(1), steps, 1)
    step_s = ev / steps
    # plain per-call form on a sample of the same problems: per stage one dop_apply and
    # torch combinations, one problem after the other
    sample = probs[:plain_sample]
    spts = sum(p[1].numel() for p in sample)

    def plain_step(A, u, sign, dt, bufs):
        plan = D.plans_for(A, u)[0]
        k, x, acc = bufs
        plan.apply(k, u)
        k.mul_(sign)
        torch.add(u, k, alpha=dt / 6, out=acc)
        torch.add(u, k, alpha=dt / 2, out=x)
        for c, w in ((dt / 2, dt / 3), (dt, dt / 3)):
            plan.apply(k, x)
            k.mul_(sign)
            acc.add_(k, alpha=w)
            torch.add(u, k, alpha=c, out=x)
        plan.apply(k, x)
        k.mul_(sign)
        torch.add(acc, k, alpha=dt / 6, out=u)

    bufs = [tuple(torch.empty_like(p[1]) for _ in range(3)) for p in sample]
    for p, b in zip(sample, bufs):
        plain_step(p[0], p[1], p[2], p[3], b)
    _sync()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for p, b in zip(sample, bufs):
            plain_step(p[0], p[1], p[2], p[3], b)
    _sync()
    plain_s = (time.perf_counter() - t0) / plain_steps
    batched_sample_s = step_s * spts / pts   # batched cost of the same points
    # algorithmic bytes of one RK4 step: stages read/write 3s, 5s, 5s, 3s bytes per point
    # (stage 0: u -> acc, s1; 1: s1, u, acc -> acc, s2; 2: s2, u, acc -> acc, s1; 3: s1, acc -> u)
    step_bytes = 16 * 4 * pts
    ach = step_bytes / step_s / 1e9
    return dict(
        roofline=dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                      bytes_per_step=step_bytes),
        problems=nprob, points=pts, rk4_step_ms=round(step_s * 1e3, 4), tile_chunk=rk.chunk,
        rhs_gpts_s=round(4 * pts / step_s / 1e9, 2),
        problem_steps_per_s=round(nprob / step_s, 1),
        solves_per_s_100_steps=round(nprob / step_s / 100, 2),
        plain_sample=dict(problems=plain_sample, points=spts, step_ms=round(plain_s * 1e3, 3),
                          problem_steps_per_s=round(plain_sample / plain_s, 1)),
        speedup_vs_plain_same_points=round(plain_s / batched_sample_s, 2),
        create_s=round(t_create, 2))


def served_requests(nreq, lam, seed=5):
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    rng = np.random.default_rng(seed)
    rel = np.cumsum(rng.exponential(1.0 / lam, nreq))
    reqs = []
    for i in range(nreq):
        r = np.random.default_rng(5000 + i)
        n = int(np.exp(r.uniform(np.log(64), np.log(2048))))
        h = 1.0 / (n + 1)
        vx, vy = r.standard_normal(n), r.standard_normal(n)
        ops = [D.CenteredDifference[1](2, 4, h, n), D.CenteredDifference[2](2, 4, h, n),
               D.UpwindDifference[1](1, 3, h, n, -vx), D.UpwindDifference[2](1, 3, h, n, -vy)]
        A = (ops[0] + ops[1] + ops[2] + ops[3]) * D.compose(*D.Dirichlet0BC(float, (n, n)))
        wsum = sum(float(np.abs(L.stencil_coefs).sum()) for L in ops[:2]) + \
            float(np.abs(np.r_[vx, vy]).max()) * sum(float(np.abs(L.stencil_coefs).sum()) for L in ops[2:])
        u0 = torch.from_numpy(r.uniform(-1, 1, (n, n)).astype(np.float32)).pin_memory()
        reqs.append((A, n, u0, 0.4 / wsum, float(rel[i])))
    return reqs


def bench_served(nreq=200, lam=2000.0, nsteps=100, budget=8 << 30):
    """Poisson-arrival method-of-lines solves: continuous batching vs the plain per-call
    form (one request at a time, per stage one dop_apply and torch combinations)."""
    import numpy as np
    import torch
    import diffeqoperators_amd as D
    reqs = served_requests(nreq, lam)
    plans = [D.plan_for_shape(A, (n, n), torch.float32) for A, n, _, _, _ in reqs]
    srv = D.Server(budget, max_active=4096, chunk=16)
    for (A, n, u0, dt, rel), plan in zip(reqs, plans):
        srv.submit(plan, u0, 1.0, dt, nsteps, rel)
    srv.run()
    t = srv.times()
    lat = t[:, 2] - t[:, 0]
    makespan = t[:, 2].max() - t[:, 0].min()
    # algorithmic bytes: 16 * 4 bytes per point per RK4 step (FP32), over the makespan (which
    # includes the arrival gaps: a lower bound on the rate while requests are resident)
    tot_bytes = sum(64.0 * n * n * nsteps for _, n, _, _, _ in reqs)
    ach = tot_bytes / makespan / 1e9
    sizes = np.array([n * n for _, n, _, _, _ in reqs], dtype=np.float64)
    cuts = np.percentile(sizes, [100 / 3, 200 / 3])
    terc = np.digitize(sizes, cuts)    # 0: smallest third of the grids, 2: largest

    def by_tercile(l):
        return {f"tercile{k}": dict(points_max=int(sizes[terc == k].max()),
                                     p50_ms=round(float(np.percentile(l[terc == k], 50)) * 1e3, 2),
                                     p99_ms=round(float(np.percentile(l[terc == k], 99)) * 1e3, 2))
                for k in range(3) if np.any(terc == k)}

    served = dict(solves_per_s=round(nreq / makespan, 1), p50_ms=round(np.percentile(lat, 50) * 1e3, 2),
                  p99_ms=round(np.percentile(lat, 99) * 1e3, 2), makespan_s=round(makespan, 3),
                  roofline=dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                                bytes_total=tot_bytes),
                  by_size=by_tercile(lat), **srv.stats())
    # plain: FIFO, one request at a time
    lat_p, t0 = [], None
    res_plain = []
    start = time.perf_counter()
    for (A, n, u0, dt, rel), plan in zip(reqs, plans):
        now = time.perf_counter() - start
        if now < rel:
            time.sleep(rel - now)
        u = u0.to("cuda", non_blocking=True)
        k, x, acc = torch.empty_like(u), torch.empty_like(u), torch.empty_like(u)
        for _ in range(nsteps):
            plan.apply(k, u)
            torch.add(u, k, alpha=dt / 6, out=acc)
            torch.add(u, k, alpha=dt / 2, out=x)
            for c, w in ((dt / 2, dt / 3), (dt, dt / 3)):
                plan.apply(k, x)
                acc.add_(k, alpha=w)
                torch.add(u, k, alpha=c, out=x)
            plan.apply(k, x)
            torch.add(acc, k, alpha=dt / 6, out=u)
        res_plain.append(u.cpu())
        lat_p.append(time.perf_counter() - start - rel)
    mk_p = time.perf_counter() - start - reqs[0][4]
    lat_p = np.array(lat_p)
    plain = dict(solves_per_s=round(nreq / mk_p, 1), p50_ms=round(np.percentile(lat_p, 50) * 1e3, 2),
                 p99_ms=round(np.percentile(lat_p, 99) * 1e3, 2), makespan_s=round(mk_p, 3),
                 by_size=by_tercile(lat_p))
    # equal accuracy: the two forms integrate the same RK4 (FP32 storage)
    diff = max(float((srv.results[i] - res_plain[i]).abs().max() / res_plain[i].abs().max())
               for i in range(nreq))
    return dict(requests=nreq, arrival_rate_per_s=lam, rk4_steps=nsteps, served=served, plain=plain,
                served_over_plain_solves=round(served["solves_per_s"] / plain["solves_per_s"], 2),
                max_rel_diff_served_vs_plain=diff)


# ------------------------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="nonuniform3d",
                    choices=["nonuniform3d", "laplacian2d", "batched", "served"])
    ap.add_argument("--no-also", action="store_true",
                    help="skip the other workloads (2-D Laplacian, batched RK4, served, ...)")
    ap.add_argument("--requests", type=int, default=200, help="served workload: requests")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the oracle legs (CPU baseline and the errors against FP64)")
    ap.add_argument("--copy-reference", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates HBM counters)")
    ap.add_argument("--types", default="f32,f64,bf16",
                    help="storage types of the 3-D workload to run (profiling runs pick one)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): there is no multi-GPU "
              "path (the operator would shard as slabs along the slowest axis, one halo "
              "exchange per apply).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    if args.copy_reference:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    base = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic"}
    if args.config == "nonuniform3d":
        r3, cpu = bench_nonuniform3d(args.steps, args.warmup, errors=not args.no_cpu,
                                     cpu_seconds=args.cpu_seconds,
                                     types=tuple(args.types.split(",")))
        head = "f32" if "f32" in r3 else next(iter(r3))
        f32 = r3[head]
        line = dict(base, value=f32["gpts_s"], unit="Gpts/s", steps=args.steps,
                    warmup=args.warmup, ms_per_step=f32["ms_per_step"], dtype=head)
        line["data"] = "synthetic (seeded dx, wind and field)"
        line["config"] = {
            "workload": "3-D non-uniform 512^3: per axis CenteredDifference{k}(2,6,dx_k,512) + "
                        "UpwindDifference{k}(1,3,dx_k,512,-v_k), Dirichlet0BC on every face, "
                        "FP32 storage and accumulation, one (sum L)*Q*u per step",
            "grid": [512, 512, 512], "storage": head, "accumulation": "f64" if head == "f64" else "f32",
            "tile_variant": f32["tile_variant"]}
        line["roofline"] = {"bound": "hbm", "achieved": f32["achieved_GBs"], "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": f32["frac"],
                            "traffic": read_traffic("apply_nonuniform3d_512_f32")}
        line["max_rel_err_vs_f64"] = f32.get("max_rel_err_vs_f64")
        line["variants"] = r3
        if cpu is not None:
            line["cpu_baseline"] = cpu
    elif args.config == "laplacian2d":
        r = bench_laplacian2d(args.steps, args.warmup)
        line = dict(base, value=round(r["value"], 3), unit="Gpts/s", steps=args.steps,
                    warmup=args.warmup, ms_per_step=round(r["ms_per_step"], 5), dtype="f64")
        line["data"] = "synthetic (seeded uniform [-1,1] field)"
        line["config"] = {"workload": "2-D Laplacian CenteredDifference{1,2}(2,4,h,4096) + "
                                      "RobinBC(1,0.5,0) on all edges, 4096x4096 FP64, one Δ*Q*u per step",
                          "grid": [4096, 4096], "storage": "f64", "accumulation": "f64",
                          "tile_variant": r["variant"]}
        line["roofline"] = r["roofline"]
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline_laplacian2d(args.cpu_seconds)
    elif args.config == "batched":
        rb = bench_batched(max(10, args.steps // 10))
        line = dict(base, value=rb["rhs_gpts_s"], unit="Gpts/s (RHS evaluations)",
                    steps=max(10, args.steps // 10), warmup=1, ms_per_step=rb["rk4_step_ms"],
                    dtype="f32")
        line["config"] = {"workload": "1024 independent 1-D/2-D Δ*Q*u problems, one RK4 step per step"}
        line["roofline"] = rb.pop("roofline")
        line["batched"] = rb
    else:
        rs = bench_served(args.requests)
        line = dict(base, value=rs["served"]["solves_per_s"], unit="solves/s", steps=args.requests,
                    warmup=0, ms_per_step=round(1e3 / max(rs["served"]["solves_per_s"], 1e-9), 3),
                    dtype="f32")
        line["config"] = {"workload": "Poisson-arrival 2-D advection-diffusion method-of-lines "
                                      "solves, 100 RK4 steps each"}
        line["roofline"] = rs["served"].pop("roofline")
        line["served"] = rs
    if not args.no_also and args.config == "nonuniform3d":
        also = line["also"] = {}
        r2 = bench_laplacian2d(args.steps, args.warmup)
        also["laplacian2d_4096_f64"] = dict(
            gpts_s=round(r2["value"], 2), ms_per_step=round(r2["ms_per_step"], 5),
            kernel_ms=round(r2["kernel_ms"], 4), roofline=r2["roofline"],
            tile_variant=r2["variant"])
        if not args.no_cpu:
            also["laplacian2d_4096_f64"]["cpu_baseline"] = cpu_baseline_laplacian2d(
                max(3.0, args.cpu_seconds / 3))
        torch.cuda.empty_cache()
        also["laplacian2d_4096_f32"] = bench_laplacian2d_f32(args.steps, args.warmup)
        torch.cuda.empty_cache()
        also["slab_512x512x70_f32"] = bench_slab()
        torch.cuda.empty_cache()
        try:
            also["widened"] = bench_widened()
        except Exception as e:   # reported, never fatal to the line
            also["widened"] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
        also["batched_rk4_1024"] = bench_batched(max(10, args.steps // 10))
        torch.cuda.empty_cache()
        also["served_advdiff"] = bench_served(args.requests)
        torch.cuda.empty_cache()
        also["copy_reference"] = bench_copy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
