#!/usr/bin/env python3
"""Throughput of the fused operator application on one MI355X.

Default workload (`--config laplacian2d`): the 2-D uniform Laplacian
CenteredDifference{1}(2,4,h,4096) + CenteredDifference{2}(2,4,h,4096) composed with
RobinBC(α=1, β=0.5, γ=0) on every edge, u a 4096×4096 FP64 field seeded uniform in [-1,1]:
one step = one Δ*Q*u (one kernel launch).  Prints ONE JSON line.

Other workloads are reported in the same line under "also" (`--also`), or selected with
`--config nonuniform3d|batched|served`.
"""
import argparse
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
    """Wall time and HIP-event time (on torch's current stream, where libdop launches)."""
    import torch
    for _ in range(warmup):
        fn()
    _sync()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    _sync()
    wall = time.perf_counter() - t0
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
    variant = plan.autotune()
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
    po.set_threads(1)
    reps, t0 = 0, time.perf_counter()
    while True:
        po.ghost_apply(ops, [bc, bc], u)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return dict(value=round(n * n * reps / el / 1e9, 4), unit="Gpts/s", cores=1, kind="port",
                sample=f"{reps} full 4096x4096 FP64 applies of the oracle (per-line loops, one "
                       f"temporary per operator), {el:.1f} s")


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


def bench_nonuniform3d(steps, warmup, n=512):
    import torch
    import diffeqoperators_amd as D
    out = {}
    for name, dt, sb in (("f64", torch.float64, 8), ("f32", torch.float32, 4), ("bf16", torch.bfloat16, 2)):
        A, u = nonuniform3d(n, dt)
        plan = D.plans_for(A, u)[0]
        variant = plan.autotune()
        du = torch.empty_like(u)
        wall, ev = time_applies(lambda: plan.apply(du, u), steps, warmup)
        pts = n ** 3
        kern = ev / steps
        ach = 2 * sb * pts / kern / 1e9
        out[name] = dict(gpts_s=round(pts / kern / 1e9, 2), kernel_ms=round(kern * 1e3, 4),
                         achieved_GBs=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4),
                         tile_variant=variant)
        del A, u, du, plan
        torch.cuda.empty_cache()
    return out


# ------------------------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="laplacian2d",
                    choices=["laplacian2d", "nonuniform3d"])
    ap.add_argument("--also", action="store_true", help="add the other workloads under 'also'")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build has no "
              "multi-GPU path; see DESIGN.md for how the operator would shard.", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    if args.config == "laplacian2d":
        r = bench_laplacian2d(args.steps, args.warmup)
        line = {
            "metric": METRIC, "value": round(r["value"], 3), "unit": "Gpts/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded uniform [-1,1] field)",
            "config": {"workload": "2-D Laplacian CenteredDifference{1,2}(2,4,h,4096) + "
                                   "RobinBC(1,0.5,0) on all edges, 4096x4096 FP64, one Δ*Q*u per step",
                       "grid": [4096, 4096], "storage": "f64", "accumulation": "f64",
                       "tile_variant": r["variant"]},
            "roofline": r["roofline"],
        }
    else:
        r3 = bench_nonuniform3d(args.steps, args.warmup)
        f32 = r3["f32"]
        line = {
            "metric": METRIC, "value": f32["gpts_s"], "unit": "Gpts/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": f32["kernel_ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "3-D non-uniform 512^3 centered(2,6)+upwind(1,3)"},
            "roofline": {"bound": "hbm", "achieved": f32["achieved_GBs"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": f32["frac"], "traffic": None},
            "variants": r3,
        }
    if args.also and args.config == "laplacian2d":
        line["also"] = {"nonuniform3d_512": bench_nonuniform3d(max(20, args.steps // 4), 3)}
    if not args.no_cpu and args.config == "laplacian2d":
        line["cpu_baseline"] = cpu_baseline_laplacian2d(args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
