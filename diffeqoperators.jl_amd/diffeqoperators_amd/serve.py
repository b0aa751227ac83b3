"""Continuous-batching method-of-lines service (dop_server_*, no counterpart in the
reference): requests are released at their arrival times, admitted FIFO while their device
state fits the HBM budget, advanced one RK4 step per iteration within their size class (each
class on its own stream, the classes side by side on the GPU) and copied back when done.
Requests may be submitted from another thread while `run` runs (`open()` keeps it waiting for
them until `close()`)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi


class Server:
    def __init__(self, hbm_budget_bytes: int, max_active: int = 1024, chunk: int = 16,
                 stream=None):
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        h = C.c_void_p()
        _abi.check(_abi.lib().dop_server_create(int(hbm_budget_bytes), int(max_active), int(chunk),
                                                C.c_void_p(stream), C.byref(h)),
                   "dop_server_create")
        self.handle = h
        self._keep = []
        self.results = {}   # request id -> pinned host tensor with the final state

    def submit(self, plan, u0: torch.Tensor, scale: float, dt: float, nsteps: int,
               release_s: float = None) -> int:
        """u0: pinned host tensor with the plan's size and dtype; returns the request id.
        The final state lands in self.results[id] (pinned host tensor).  release_s: seconds
        after the run starts (None: now, or the start of the next run)."""
        if release_s is None:
            release_s = max(self.now(), 0.0)
        from .apply import storage_of
        if not isinstance(u0, torch.Tensor) or u0.device.type != "cpu" or not u0.is_pinned():
            raise ValueError("u0 must be a pinned host tensor")
        if not u0.is_contiguous():
            raise ValueError("u0 must be contiguous")
        if storage_of(u0) != plan.storage:
            raise TypeError(f"u0 has dtype {u0.dtype}, the plan's storage type differs")
        if u0.numel() != plan.in_elems:
            raise ValueError(f"u0 has {u0.numel()} elements, the plan needs {plan.in_elems}")
        out = torch.empty_like(u0).pin_memory()
        rid = C.c_int64()
        _abi.check(_abi.lib().dop_server_submit(
            self.handle, plan.handle, float(scale), float(dt), int(nsteps),
            C.c_void_p(u0.data_ptr()), C.c_void_p(out.data_ptr()), float(release_s),
            C.byref(rid)), "dop_server_submit")
        self._keep.append((plan, u0, out))
        self.results[rid.value] = out
        return rid.value

    def run(self):
        """Serves until every submitted request is done and the server is not open (the call
        releases the GIL: other threads may submit meanwhile)."""
        _abi.check(_abi.lib().dop_server_run(self.handle), "dop_server_run")

    def open(self):
        """Keep run() waiting for submissions (until close())."""
        _abi.check(_abi.lib().dop_server_set_open(self.handle, 1), "dop_server_set_open")

    def close(self):
        _abi.check(_abi.lib().dop_server_set_open(self.handle, 0), "dop_server_set_open")

    def now(self):
        """Seconds since the running run() started (-1 when not running)."""
        return _abi.lib().dop_server_now(self.handle)

    def times(self):
        """(release, admitted, completed) seconds per request, relative to the run start."""
        n = len(self.results)
        out = np.zeros((n, 3))
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        for i in range(n):
            _abi.check(_abi.lib().dop_server_times(self.handle, i, C.byref(a), C.byref(b),
                                                   C.byref(c)), "dop_server_times")
            out[i] = (a.value, b.value, c.value)
        return out

    def stats(self):
        pa, pb, it = C.c_int(), C.c_int64(), C.c_int64()
        _abi.check(_abi.lib().dop_server_stats(self.handle, C.byref(pa), C.byref(pb), C.byref(it)),
                   "dop_server_stats")
        return dict(peak_active=pa.value, peak_bytes=pb.value, iterations=it.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_server_destroy(h)
            except Exception:
                pass
            self.handle = None
