"""Adaptive time stepping of a method-of-lines problem on the device.

The reference's tests integrate `step = (u, p, t) -> A*bc*u` with OrdinaryDiffEq's Tsit5
(heat_eqn.jl:4-60).  `Tsit5(A, u, tspan)` runs the same kind of solve entirely on the GPU
(libdop's dop_tsit5_*): Tsitouras' 5(4) pair, the scaled RMS error norm with abstol / reltol
(OrdinaryDiffEq's defaults 1e-6 / 1e-3), the DOPRI5 step controller, one hipGraph replay per
attempted step.

    I = Tsit5(Δ * bc, u, (0.0, 1.0))     # u: float64 / float32 tensor, integrated in place
    I.solve()                             # or I.advance_to(t) for intermediate times
    I.t, I.naccept, I.nreject
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .apply import _ghost_terms, ghost_plans

__all__ = ["Tsit5"]


class Tsit5:
    """du/dt = scale * A u on the device, in place in `u`.  A: L*Q or a sum of them (one
    fused plan)."""

    def __init__(self, A, u, tspan, *, scale=1.0, dt=None, abstol=1e-6, reltol=1e-3):
        if not isinstance(u, torch.Tensor) or u.device.type != "cuda":
            raise TypeError("u must be a torch tensor on the GPU (no CPU path)")
        if u.dtype not in (torch.float64, torch.float32) or not u.is_contiguous():
            raise TypeError("u must be a contiguous float64 or float32 tensor")
        terms = _ghost_terms(A)
        if terms is None:
            raise TypeError("Tsit5 needs L*Q (a GhostDerivativeOperator) or a sum of them")
        plans = ghost_plans(A, terms, u)
        if len(plans) != 1:
            raise ValueError("the operators' BCs must agree on every axis (one fused plan)")
        self._plan = plans[0]
        self.u = u
        self.t0, self.t1 = float(tspan[0]), float(tspan[1])
        h = C.c_void_p()
        with torch.cuda.device(u.device):
            _abi.check(_abi.lib().dop_tsit5_create(
                self._plan.handle, float(scale), C.c_void_p(u.data_ptr()), self.t0, self.t1,
                float(dt) if dt else 0.0, float(abstol), float(reltol),
                C.c_void_p(self._stream()), C.byref(h)), "dop_tsit5_create")
        self.handle = h

    def _stream(self):
        return torch.cuda.current_stream(self.u.device).cuda_stream

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_tsit5_destroy(h)
            except Exception:
                pass
            self.handle = None

    def run(self, max_attempts=10 ** 6, check_every=64):
        """Attempts steps until the end time (or max_attempts); returns self."""
        _abi.check(_abi.lib().dop_tsit5_run(self.handle, int(max_attempts), int(check_every),
                                            C.c_void_p(self._stream())), "dop_tsit5_run")
        return self

    def solve(self, max_attempts=10 ** 6):
        self.run(max_attempts)
        if self.status()["done"] != 1:
            raise RuntimeError(f"Tsit5 stopped at t = {self.t} before {self.t1}: {self.status()}")
        return self.u

    def advance_to(self, t, max_attempts=10 ** 6):
        """Moves the end time to t and integrates up to it (u then holds u(t))."""
        _abi.check(_abi.lib().dop_tsit5_set_end(self.handle, float(t), C.c_void_p(self._stream())),
                   "dop_tsit5_set_end")
        self.t1 = float(t)
        return self.solve(max_attempts)

    def status(self):
        t, dt = C.c_double(), C.c_double()
        na, nr = C.c_int64(), C.c_int64()
        done = C.c_int()
        _abi.check(_abi.lib().dop_tsit5_status(self.handle, C.byref(t), C.byref(dt), C.byref(na),
                                               C.byref(nr), C.byref(done),
                                               C.c_void_p(self._stream())), "dop_tsit5_status")
        return {"t": t.value, "dt": dt.value, "naccept": na.value, "nreject": nr.value,
                "done": done.value}

    @property
    def t(self):
        return self.status()["t"]

    @property
    def naccept(self):
        return self.status()["naccept"]

    @property
    def nreject(self):
        return self.status()["nreject"]
