"""nonlinear_diffusion! / nonlinear_diffusion (derivative_operator.jl:34-77) on the device.

    du = sum_{l=0..s} C(s, l) (D^{l+f} q) .* (D^{s-l} p),   D^0 = the interior q[2:n+1]

(the Leibniz expansion of d^s/dx^s (p d^f q/dx^f)), with D^k = CenteredDifference(k,
approx_order, dx, nknots) on the padded vectors p and q (length nknots + 2).  Every term of a
point is evaluated in one libdop launch (dop_nonlinear_diffusion): on uniform grids one thread
takes four points and reads their windows of p and q once; otherwise a thread per point.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .apply import Plan, _check_device_tensor, make_desc, storage_of
from .operators import CenteredDifference

__all__ = ["nonlinear_diffusion_", "nonlinear_diffusion"]

_CACHE = {}


def _plans(s, f, approx_order, dx, nknots, storage, device):
    dkey = float(dx) if np.ndim(dx) == 0 else np.asarray(dx, dtype=np.float64).tobytes()
    key = (s, f, approx_order, dkey, nknots, storage, str(device))
    if key not in _CACHE:
        def plan(k):
            if k == 0:
                return None
            L = CenteredDifference(k, approx_order, dx, nknots)
            d, keep = make_desc([L], (nknots,), storage, input_kind=_abi.INPUT_PADDED,
                                semantics=_abi.SEM_LINES, padded=(1, 0, 0))
            with torch.cuda.device(device):
                return Plan(d, keep)
        _CACHE[key] = ([plan(f + l) for l in range(s + 1)], [plan(s - l) for l in range(s + 1)])
    return _CACHE[key]


def nonlinear_diffusion_(du, second_differential_order, first_differential_order, approx_order,
                         p, q, dx, nknots):
    """nonlinear_diffusion!(du, second_differential_order, first_differential_order,
    approx_order, p, q, dx, nknots) (derivative_operator.jl:34-66)."""
    s, f = int(second_differential_order), int(first_differential_order)
    if approx_order <= 1:
        raise AssertionError("approximation_order must be greater than 1.")
    if s < 1 or f < 0 or s > 8:
        raise ValueError("need 1 <= second_differential_order <= 8, first_differential_order >= 0")
    for t, name in ((du, "du"), (p, "p"), (q, "q")):
        _check_device_tensor(t, name)
    if p.dim() != 1 or q.dim() != 1 or du.dim() != 1:
        raise AssertionError("nonlinear_diffusion! works on vectors")
    if p.numel() != nknots + 2 or q.numel() != nknots + 2 or du.numel() != nknots:
        raise AssertionError("p and q must have nknots + 2 entries, du nknots")
    if not (p.dtype == q.dtype == du.dtype):
        raise TypeError("du, p and q must share one dtype")
    storage = storage_of(q)
    pq, pp = _plans(s, f, int(approx_order), dx, int(nknots), storage, q.device)
    vp = C.c_void_p
    aq = (vp * (s + 1))(*[x.handle if x is not None else None for x in pq])
    ap = (vp * (s + 1))(*[x.handle if x is not None else None for x in pp])
    _abi.check(_abi.lib().dop_nonlinear_diffusion(
        aq, ap, s + 1, vp(du.data_ptr()), vp(p.data_ptr()), vp(q.data_ptr()),
        vp(torch.cuda.current_stream(q.device).cuda_stream)), "dop_nonlinear_diffusion")
    return du


def nonlinear_diffusion(second_differential_order, first_differential_order, approx_order, p, q,
                        dx, nknots):
    """The out-of-place form (derivative_operator.jl:68-77)."""
    _check_device_tensor(q, "q")
    du = torch.empty(q.numel() - 2, dtype=q.dtype, device=q.device)
    return nonlinear_diffusion_(du, second_differential_order, first_differential_order,
                                approx_order, p, q, dx, nknots)
