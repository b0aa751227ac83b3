"""Many independent problems per launch (no counterpart in the reference, which applies
one operator per call): a Batch evaluates the RHS of every problem in one persistent
kernel; RK4 advances them all with graph-captured classical RK4 steps."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi


def _single_plan(A, u):
    from .apply import plans_for
    plans = plans_for(A, u)
    if len(plans) != 1:
        raise ValueError("batched problems need one fused plan per problem (one BC per axis)")
    return plans[0]


def _check_buffer(t, plan, n_elems, name, i):
    """A device buffer the library reads or writes for problem i: a contiguous GPU tensor of
    the plan's storage type and element count (the C side trusts the pointer)."""
    from .apply import _check_device_tensor, storage_of
    _check_device_tensor(t, f"{name}[{i}]")
    if storage_of(t) != plan.storage:
        raise TypeError(f"{name}[{i}] has dtype {t.dtype}, the plan's storage type differs")
    if t.numel() != n_elems:
        raise ValueError(f"{name}[{i}] has {t.numel()} elements, the plan needs {n_elems}")


class Batch:
    """du_i = A_i * u_i for all i in one launch (dop_batch_create / dop_batch_apply)."""

    def __init__(self, ops, us, dus):
        if not (len(ops) == len(us) == len(dus)) or not ops:
            raise ValueError("need equal, non-empty lists of operators, inputs and outputs")
        from .apply import _check_device_tensor
        for i, (u, du) in enumerate(zip(us, dus)):
            _check_device_tensor(u, f"us[{i}]")
            _check_device_tensor(du, f"dus[{i}]")
        self.plans = [_single_plan(A, u) for A, u in zip(ops, us)]
        for i, (p, u, du) in enumerate(zip(self.plans, us, dus)):
            _check_buffer(u, p, p.in_elems, "us", i)
            _check_buffer(du, p, p.out_elems, "dus", i)
            if du.data_ptr() == u.data_ptr():
                raise ValueError(f"dus[{i}] must not alias us[{i}]")
        self.us, self.dus = list(us), list(dus)
        n = len(ops)
        P = (C.c_void_p * n)(*[p.handle.value for p in self.plans])
        D = (C.c_void_p * n)(*[d.data_ptr() for d in dus])
        U = (C.c_void_p * n)(*[u.data_ptr() for u in us])
        h = C.c_void_p()
        _abi.check(_abi.lib().dop_batch_create(n, P, D, U, C.byref(h)), "dop_batch_create")
        self.handle = h

    def apply(self, accumulate=False, stream=None):
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _abi.check(_abi.lib().dop_batch_apply(self.handle, int(bool(accumulate)),
                                              C.c_void_p(stream)), "dop_batch_apply")

    def autotune(self, stream=None):
        """Pick the row-chunk height of the 2-D tiles by timing (overwrites the outputs)."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        c = C.c_int()
        _abi.check(_abi.lib().dop_batch_autotune(self.handle, C.c_void_p(stream), C.byref(c)),
                   "dop_batch_autotune")
        return c.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_batch_destroy(h)
            except Exception:
                pass
            self.handle = None


class RK4:
    """du_i/dt = scale_i * A_i u_i, classical RK4 with step dt_i, updating u_i in place;
    the four stage launches of one step are a captured hipGraph (dop_rk4_*)."""

    def __init__(self, ops, us, scale, dt, stream=None):
        n = len(ops)
        if not (n == len(us) == len(scale) == len(dt)) or n == 0:
            raise ValueError("need equal, non-empty lists")
        from .apply import _check_device_tensor
        for i, u in enumerate(us):
            _check_device_tensor(u, f"us[{i}]")
        self.plans = [_single_plan(A, u) for A, u in zip(ops, us)]
        for i, (p, u) in enumerate(zip(self.plans, us)):
            _check_buffer(u, p, p.in_elems, "us", i)
        self.us = list(us)
        self._scale = np.ascontiguousarray(scale, dtype=np.float64)
        self._dt = np.ascontiguousarray(dt, dtype=np.float64)
        P = (C.c_void_p * n)(*[p.handle.value for p in self.plans])
        U = (C.c_void_p * n)(*[u.data_ptr() for u in us])
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        h = C.c_void_p()
        _abi.check(_abi.lib().dop_rk4_create(
            n, P, self._scale.ctypes.data_as(_abi._dp), self._dt.ctypes.data_as(_abi._dp), U,
            C.c_void_p(stream), C.byref(h)), "dop_rk4_create")
        self.handle = h

    @property
    def chunk(self):
        """Row-chunk height of the 2-D tiles chosen when the integrator was created."""
        return _abi.lib().dop_rk4_chunk(self.handle)

    def step(self, nsteps=1, stream=None):
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _abi.check(_abi.lib().dop_rk4_step(self.handle, int(nsteps), C.c_void_p(stream)),
                   "dop_rk4_step")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_rk4_destroy(h)
            except Exception:
                pass
            self.handle = None
