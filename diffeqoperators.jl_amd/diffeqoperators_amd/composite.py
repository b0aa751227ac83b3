"""Lazy compositions: L*Q (GhostDerivativeOperator, ghost_derivative_operator.jl:1-37) and
L1 + L2 (DiffEqOperatorCombination, composite_operators.jl:15-84)."""
from __future__ import annotations

import numbers


class GhostDerivativeOperator:
    """L * Q: applying it pads u with Q's ghosts, then applies L (ghost_derivative_operator.jl:
    15-37).  On the GPU both steps are one fused kernel."""

    __array_ufunc__ = None

    def __init__(self, L, Q):
        self.L, self.Q = L, Q
        self._plans = {}

    def __mul__(self, u):
        from .apply import ghost_times
        return ghost_times(self, u)

    def __rmul__(self, c):
        # c * (L*Q) = (c*L)*Q  (:65-76)
        return GhostDerivativeOperator(c * self.L, self.Q)

    def __add__(self, other):
        return DiffEqOperatorCombination.of(self, other)

    def __neg__(self):
        return -1.0 * self

    def __sub__(self, other):
        return self + (-other)

    @property
    def size(self):
        return (self.L.len, self.L.len)


class DiffEqOperatorCombination:
    """A lazy sum of operators; `*` and `mul!` evaluate all of them in one fused pass."""

    __array_ufunc__ = None

    def __init__(self, ops):
        self.ops = tuple(ops)
        self._plans = {}

    @staticmethod
    def of(a, b):
        la = a.ops if isinstance(a, DiffEqOperatorCombination) else (a,)
        lb = b.ops if isinstance(b, DiffEqOperatorCombination) else (b,)
        return DiffEqOperatorCombination(la + lb)

    def __add__(self, other):
        return DiffEqOperatorCombination.of(self, other)

    def __radd__(self, other):
        return DiffEqOperatorCombination.of(other, self)

    def __neg__(self):
        return DiffEqOperatorCombination(tuple(-op for op in self.ops))

    def __sub__(self, other):
        return self + (-other)

    def __rmul__(self, c):
        if isinstance(c, numbers.Number):
            return DiffEqOperatorCombination(tuple(c * op for op in self.ops))
        return NotImplemented

    def __mul__(self, other):
        from .bcs import BoundaryPadded, is_bc
        if isinstance(other, BoundaryPadded):
            return (self * other.Q) * other.u
        if is_bc(other):
            # *(L::AbstractDiffEqCompositeOperator, Q) = sum(op * Q) (ghost_derivative_operator.jl:11-13)
            return DiffEqOperatorCombination(tuple(op * other for op in self.ops))
        from .apply import combination_times
        return combination_times(self, other)
