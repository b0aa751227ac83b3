"""Derivative operators: the host-side constructors of DiffEqOperators.jl, FP64.

CenteredDifference / UpwindDifference build the same tables as
src/derivative_operators/derivative_operator.jl (uniform :81-135 and :454-512,
non-uniform :137-205 and :515-639); applying them runs on the GPU through libdop.so.

Axis convention: Julia's `CenteredDifference{N}` differentiates along array dimension N of
a column-major array.  Here arrays are C-ordered torch tensors whose shape is the Julia
size reversed, so axis N is tensor dimension -N.  `CenteredDifference[N](...)` mirrors
`CenteredDifference{N}(...)`; without [N] the axis is 1.
"""
from __future__ import annotations

import numbers

import numpy as np

from . import _abi
from .fornberg import calculate_weights


def _cumsum(v):
    out, acc = [], 0.0
    for x in v:
        acc += float(x)
        out.append(acc)
    return out


class DerivativeOperator:
    """The fields of the reference's DerivativeOperator (derivative_operator.jl:14-29),
    plus its axis N (type parameter) and Wind flag."""

    __array_ufunc__ = None   # `vector * L` reaches __rmul__ instead of numpy broadcasting

    def __init__(self, *, N, wind, derivative_order, approximation_order, dx, len,
                 stencil_length, stencil_coefs, boundary_stencil_length,
                 boundary_point_count, low_boundary_coefs, high_boundary_coefs, offside,
                 coefficients, coeff_func):
        self.N = int(N)
        self.wind = bool(wind)
        self.derivative_order = int(derivative_order)
        self.approximation_order = int(approximation_order)
        self.dx = dx
        self.len = int(len)
        self.stencil_length = int(stencil_length)
        self.stencil_coefs = np.ascontiguousarray(stencil_coefs, dtype=np.float64)
        self.boundary_stencil_length = int(boundary_stencil_length)
        self.boundary_point_count = int(boundary_point_count)
        self.low_boundary_coefs = np.ascontiguousarray(low_boundary_coefs, dtype=np.float64)
        self.high_boundary_coefs = np.ascontiguousarray(high_boundary_coefs, dtype=np.float64)
        self.offside = int(offside)
        self.coefficients = np.ascontiguousarray(coefficients, dtype=np.float64)
        self.coeff_func = coeff_func

    @property
    def uniform_grid(self) -> bool:
        return np.ndim(self.dx) == 0

    @property
    def coeff_kind(self) -> int:
        """The `coeff_func isa Number` branch of the non-uniform upwind interior kernel
        (convolutions.jl:234-253)."""
        f = self.coeff_func
        if isinstance(f, numbers.Number) and not isinstance(f, bool):
            return _abi.COEFF_NUM_POS if f >= 0 else _abi.COEFF_NUM_NEG
        return _abi.COEFF_VECTOR

    def use_winding(self) -> bool:
        return self.wind

    def diff_axis(self) -> int:
        return self.N

    def _replace_coefficients(self, coefficients):
        return DerivativeOperator(
            N=self.N, wind=self.wind, derivative_order=self.derivative_order,
            approximation_order=self.approximation_order, dx=self.dx, len=self.len,
            stencil_length=self.stencil_length, stencil_coefs=self.stencil_coefs,
            boundary_stencil_length=self.boundary_stencil_length,
            boundary_point_count=self.boundary_point_count,
            low_boundary_coefs=self.low_boundary_coefs,
            high_boundary_coefs=self.high_boundary_coefs, offside=self.offside,
            coefficients=coefficients, coeff_func=self.coeff_func)

    # c * A (derivative_operator_functions.jl:165-197): folds c into the coefficients
    def __rmul__(self, c):
        if isinstance(c, numbers.Number):
            return self._replace_coefficients(float(c) * self.coefficients)
        c = np.asarray(c, dtype=np.float64)
        if c.ndim == 1:
            if len(c) != self.len:
                raise ValueError(
                    f"DimensionMismatch: length of c ({len(c)}) must match length of A ({self.len})")
            return self._replace_coefficients(c * self.coefficients)
        return NotImplemented

    def __neg__(self):
        return -1.0 * self

    def __add__(self, other):
        from .composite import DiffEqOperatorCombination
        return DiffEqOperatorCombination.of(self, other)

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        from .bcs import BoundaryPadded, is_bc
        from .composite import GhostDerivativeOperator
        if is_bc(other):
            return GhostDerivativeOperator(self, other)   # ghost_derivative_operator.jl:7-9
        if isinstance(other, BoundaryPadded):
            # L * (Q*u): what mul!(out, L*Q, u) evaluates (ghost_derivative_operator.jl:15-24)
            return GhostDerivativeOperator(self, other.Q) * other.u
        from .apply import derivative_times_array
        return derivative_times_array(self, other)

    def c_struct(self, axis0: int) -> _abi.DopOperator:
        def ptr(a):
            return a.ctypes.data_as(_abi._dp) if a.size else None
        return _abi.DopOperator(
            axis0, int(self.wind), self.derivative_order, self.approximation_order, self.len,
            self.stencil_length, self.boundary_stencil_length, self.boundary_point_count,
            self.offside, int(self.uniform_grid), self.coeff_kind, 0,
            ptr(self.stencil_coefs), ptr(self.low_boundary_coefs),
            ptr(self.high_boundary_coefs), ptr(self.coefficients))

    def __repr__(self):
        kind = "UpwindDifference" if self.wind else "CenteredDifference"
        grid = "uniform" if self.uniform_grid else "non-uniform"
        return (f"{kind}{{{self.N}}}(d={self.derivative_order}, a={self.approximation_order}, "
                f"len={self.len}, {grid})")


def _coefficients(coeff_func, n):
    """compute_coeffs!(coeff_func, zeros(len)) (coefficient_functions.jl:69-108)."""
    cur = np.zeros(n)
    if isinstance(coeff_func, numbers.Number):
        return cur + float(coeff_func)
    if callable(coeff_func):
        try:
            v = np.asarray(coeff_func(cur.copy()), dtype=np.float64)
            if v.shape == (n,):
                return v
        except Exception:
            pass
        return np.array([float(coeff_func(x)) for x in cur])
    v = np.asarray(coeff_func, dtype=np.float64)
    if v.shape != (n,):
        raise ValueError("coefficient vector length must equal len")
    return cur + v


def _centered(N, derivative_order, approximation_order, dx, len, coeff_func=1):
    d, a, n = int(derivative_order), int(approximation_order), int(len)
    SL = d + a - 1 + (d + a) % 2
    BSL = d + a
    BPC = SL // 2 - 1
    if np.ndim(dx) == 0:
        # derivative_operator.jl:81-135
        if not a > 1:
            raise AssertionError("approximation_order must be greater than 1.")
        scale = 1.0 / (float(dx) ** d)
        half = SL // 2
        stencil = [scale * w for w in calculate_weights(d, 0.0, range(-half, half + 1))]
        low = [[scale * w for w in calculate_weights(d, float(x0), range(BSL))]
               for x0 in range(1, BPC + 1)]
        sgn = 1.0 if d % 2 == 0 else -1.0
        high = [[low[BPC - 1 - i][BSL - 1 - k] * sgn for k in range(BSL)] for i in range(BPC)]
        dxv = float(dx)
    else:
        # derivative_operator.jl:137-205 -- dx has len+1 spacings
        dxv = np.asarray(dx, dtype=np.float64)
        if dxv.shape != (n + 1,):
            raise ValueError("non-uniform dx must have len+1 entries")
        half = SL // 2
        stencil = []
        for i in range(BPC + 2, n + 2 - BPC):   # padded point index (1-based)
            sx = [0.0] * SL
            for idx in range(1, half + 1):
                sx[half + idx] = sx[half + idx - 1] + dxv[i + idx - 2]
                sx[half - idx] = sx[half - idx + 1] - dxv[i - idx - 1]
            stencil.append(calculate_weights(d, 0.0, sx))
        lx = [0.0] + _cumsum(dxv[: BSL - 1])
        low = [calculate_weights(d, lx[i], lx) for i in range(1, BPC + 1)]
        hx = _cumsum(dxv[n + 1 - BSL:])
        high = [calculate_weights(d, hx[BSL - 1 - i], hx) for i in range(BPC, 0, -1)]
    return DerivativeOperator(
        N=N, wind=False, derivative_order=d, approximation_order=a, dx=dxv, len=n,
        stencil_length=SL, stencil_coefs=np.array(stencil, dtype=np.float64),
        boundary_stencil_length=BSL, boundary_point_count=BPC,
        low_boundary_coefs=np.array(low, dtype=np.float64).reshape(BPC, BSL),
        high_boundary_coefs=np.array(high, dtype=np.float64).reshape(BPC, BSL), offside=0,
        coefficients=_coefficients(coeff_func, n), coeff_func=coeff_func)


def _upwind(N, derivative_order, approximation_order, dx, len, coeff_func=1, *, offside=0):
    d, a, n, off = int(derivative_order), int(approximation_order), int(len), int(offside)
    if not off > -1:
        raise AssertionError("Number of offside points should be non-negative")
    if not off <= (d + a - 1) // 2:
        raise AssertionError("Number of offside points should not exceed the primary wind points")
    SL = BSL = d + a
    BPC = BSL - 2 - off
    if np.ndim(dx) == 0:
        # derivative_operator.jl:454-512
        dxv = float(dx)
        scale = 1.0 / (dxv ** d)
        stencil = [scale * w for w in calculate_weights(d, 0.0, [float(k - off) for k in range(SL)])]
        low = [[scale * w for w in calculate_weights(d, float(x0), range(BSL))]
               for x0 in range(1, BPC + 1)]
        hscale = (-1.0 / dxv) ** d
        hx = [-float(k) for k in range(BSL)]
        high = [[hscale * w for w in calculate_weights(d, -float(m), hx)]
                for m in range(1, BSL - 1)]
        low_arr = np.array(low, dtype=np.float64).reshape(BPC, BSL)
        high_arr = np.array(high, dtype=np.float64).reshape(BPC + off, BSL)
        st_arr = np.array(stencil, dtype=np.float64)
    else:
        # derivative_operator.jl:515-639 -- x = [0, cumsum(dx)...] (len+2 points)
        dxv = np.asarray(dx, dtype=np.float64)
        if dxv.shape != (n + 1,):
            raise ValueError("non-uniform dx must have len+1 entries")
        x = [0.0] + _cumsum(dxv)

        def X(k):   # 1-based
            return x[k - 1]

        def W(i, first, cnt):
            return calculate_weights(d, X(i + 1), x[first - 1:first - 1 + cnt])

        if off == 0:
            up_low = [W(i, i + 1, BSL) for i in range(1, BPC + 1)]
        else:
            up_low = ([W(i, 1, BSL) for i in range(1, off)] +
                      [W(i, i + 1 - off, BSL) for i in range(off, BPC + 1)])
        dn_low = [W(i, 1, BSL) for i in range(1, BPC + 1)]
        up_st = [W(i, i + 1 - off, SL) for i in range(BPC + 1, n - BPC + 1)]
        dn_st = [W(i, i - SL + 2 + off, SL) for i in range(BPC + 1, n - BPC + 1)]
        up_high = [W(i, n - BSL + 3, BSL) for i in range(n - BPC + 1 - off, n + 1)]
        if off == 0:
            dn_high = [W(i, i - SL + 2, BSL) for i in range(n - BPC + 1, n + 1)]
        elif off == 1:
            dn_high = [W(i, i - SL + 2 + off, BSL) for i in range(n - BPC + 1 - off, n - off + 2)]
        else:
            dn_high = ([W(i, i - SL + 2 + off, BSL) for i in range(n - BPC + 1 - off, n - off + 2)] +
                       [W(i, n - SL + 3, BSL) for i in range(n - off + 2, n + 1)])
        nin = n - 2 * BPC
        st_arr = np.array(up_st + dn_st, dtype=np.float64).reshape(2, nin, SL)
        low_arr = np.array(up_low + dn_low, dtype=np.float64).reshape(2, BPC, BSL)
        high_arr = np.array(up_high + dn_high, dtype=np.float64).reshape(2, BPC + off, BSL)
    return DerivativeOperator(
        N=N, wind=True, derivative_order=d, approximation_order=a, dx=dxv, len=n,
        stencil_length=SL, stencil_coefs=st_arr, boundary_stencil_length=BSL,
        boundary_point_count=BPC, low_boundary_coefs=low_arr, high_boundary_coefs=high_arr,
        offside=off, coefficients=_coefficients(coeff_func, n), coeff_func=coeff_func)


class _AxisConstructor:
    """`Ctor(...)` builds along axis 1, `Ctor[N](...)` along axis N (Julia's Ctor{N})."""

    def __init__(self, fn, name, N=1):
        self._fn, self._name, self._N = fn, name, N

    def __getitem__(self, N):
        if not isinstance(N, int) or N < 1:
            raise TypeError("axis must be a positive integer")
        return _AxisConstructor(self._fn, self._name, N)

    def __call__(self, *args, **kwargs):
        return self._fn(self._N, *args, **kwargs)

    def __repr__(self):
        return f"{self._name}{{{self._N}}}"


CenteredDifference = _AxisConstructor(_centered, "CenteredDifference")
UpwindDifference = _AxisConstructor(_upwind, "UpwindDifference")
