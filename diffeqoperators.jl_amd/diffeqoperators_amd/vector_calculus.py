"""Gradient, Divergence and Curl (vector_calculus_operators.jl:5-88) on the device.

Each is a tuple of first-derivative CenteredDifference operators, one per axis; applying one
runs one fused libdop plan per derivative on the padded input (vector_calculus_convolutions.jl):
component k of `Gradient * M` is L_k along axis k, `Divergence * M` sums L_k over the
components, `Curl * M` forms (d2 u3 - d3 u2, d3 u1 - d1 u3, d1 u2 - d2 u1) with the negative
terms as plans of the negated operators.  For 2-D and 3-D arrays the reference's kernels index
the transverse axes of the padded input at 1:n (vector_calculus_convolutions.jl:188-838), which
the plans reproduce (transverse_view); 1-D has no transverse axis.

Layout: Julia's trailing component axis is the leading tensor dimension (Julia size
(n1, ..., nN, N) = tensor shape (N, nN, ..., n1)), so each component is contiguous.
"""
from __future__ import annotations

import builtins
import numbers

import torch

from .apply import _check_device_tensor, _padded_plan
from .operators import CenteredDifference

__all__ = ["Gradient", "Divergence", "Curl", "GradientOperator", "DivergenceOperator",
           "CurlOperator"]


def _ops(approximation_order, dx, length, coeff_func):
    n = builtins.len(length)
    if builtins.len(dx) != n:
        raise ValueError("dx needs one entry per dimension")
    return tuple(CenteredDifference[k + 1](1, approximation_order, dx[k], length[k], coeff_func)
                 for k in range(n))


class _VectorOperator:
    __array_ufunc__ = None

    def __init__(self, ops):
        self.ops = tuple(ops)
        self._plans = {}

    def __rmul__(self, c):
        if isinstance(c, numbers.Number):
            return type(self)(tuple(c * L for L in self.ops))
        return NotImplemented

    def _apply(self, L, du, M, accumulate=False):
        tv = 1 if M.dim() in (2, 3) else 0
        _padded_plan(self, [L], du, M, transverse_view=tv).apply(du, M, accumulate=accumulate)


class GradientOperator(_VectorOperator):
    def __mul__(self, M):
        _check_device_tensor(M, "M")
        N = builtins.len(self.ops)
        if M.dim() != N:
            raise AssertionError(f"Gradient of {N} dimensions applied to a {M.dim()}-D array")
        out = torch.empty((N,) + tuple(s - 2 for s in M.shape), dtype=M.dtype, device=M.device)
        for k, L in enumerate(self.ops):
            self._apply(L, out[k], M)
        return out


class DivergenceOperator(_VectorOperator):
    def __mul__(self, M):
        _check_device_tensor(M, "M")
        N = builtins.len(self.ops)
        if M.dim() != N + 1 or M.shape[0] != N:
            raise AssertionError(f"Divergence of {N} dimensions needs an array of shape "
                                 f"({N}, padded...)")
        out = torch.empty(tuple(s - 2 for s in M.shape[1:]), dtype=M.dtype, device=M.device)
        for k, L in enumerate(self.ops):
            self._apply(L, out, M[k], accumulate=k > 0)
        return out


class CurlOperator(_VectorOperator):
    def __init__(self, ops):
        super().__init__(ops)
        self._neg = tuple(-1.0 * L for L in self.ops)

    def __mul__(self, M):
        _check_device_tensor(M, "M")
        if M.dim() != 4 or M.shape[0] != 3:
            raise AssertionError("Curl needs an array of shape (3, n+2, n+2, n+2)")
        out = torch.empty((3,) + tuple(s - 2 for s in M.shape[1:]), dtype=M.dtype,
                          device=M.device)
        # (d2 u3 - d3 u2, d3 u1 - d1 u3, d1 u2 - d2 u1), d_k along Julia axis k (0-based k-1)
        for c, (kp, up, km, um) in enumerate(((1, 2, 2, 1), (2, 0, 0, 2), (0, 1, 1, 0))):
            self._apply(self.ops[kp], out[c], M[up])
            self._apply(self._neg[km], out[c], M[um], accumulate=True)
        return out


def Gradient(approximation_order, dx, len, coeff_func=1):   # noqa: A002 (the reference's name)
    """Gradient(approximation_order, dx, len; coeff_func) (vector_calculus_operators.jl:9-17)."""
    return GradientOperator(_ops(approximation_order, dx, len, coeff_func))


def Divergence(approximation_order, dx, len, coeff_func=1):   # noqa: A002
    """Divergence(approximation_order, dx, len; coeff_func) (vector_calculus_operators.jl:40-48)."""
    return DivergenceOperator(_ops(approximation_order, dx, len, coeff_func))


def Curl(approximation_order, dx, len, coeff_func=1):   # noqa: A002
    """Curl(approximation_order, dx, len; coeff_func) (vector_calculus_operators.jl:25-34): 3-D,
    equal extents."""
    if builtins.len(len) != 3 or not (len[0] == len[1] == len[2]):
        raise AssertionError("All dimensions must have equal no. of grid points")
    return CurlOperator(_ops(approximation_order, dx, len, coeff_func))
