"""Boundary conditions of the accelerated path: the affine (Robin-family) BCs of
src/derivative_operators/bc_operators.jl and their N-D extensions from
multi_dim_bc_operators.jl.

An affine BC pads a line u with ghosts l = a_l . u[1:k] + b_l and r = a_r . u[end-k+1:end]
+ b_r (bc_operators.jl:188-191).  Dirichlet is Robin with beta = 0 evaluated through IEEE
division by zero (a = +-0.0), exactly as in the reference (:176-178).  PeriodicBC pads a line
with u[end], u[1] (:192); extended over an N-D array by a directional or composed BC whose
atoms are all periodic, its ghost planes are the first and last planes of u instead
(multi_dim_bc_operators.jl:221-252).
"""
from __future__ import annotations

import numbers

import numpy as np

from .fornberg import calculate_weights


class AffineBC:
    """RobinBC / DirichletBC / NeumannBC (all Robin): a_l, b_l, a_r, b_r."""

    def __init__(self, a_l, b_l, a_r, b_r):
        self.a_l = np.asarray(a_l, dtype=np.float64)
        self.b_l = float(b_l)
        self.a_r = np.asarray(a_r, dtype=np.float64)
        self.b_r = float(b_r)

    def key(self):
        return (tuple(self.a_l.tolist()), self.b_l, tuple(self.a_r.tolist()), self.b_r)

    def __repr__(self):
        return f"AffineBC(a_l={self.a_l.tolist()}, b_l={self.b_l}, a_r={self.a_r.tolist()}, b_r={self.b_r})"

    def __mul__(self, u):
        # Q * u (bc_operators.jl:188-192, multi_dim_bc_operators.jl:210-252): lazy padding
        return _bc_times(self, u)


class PeriodicAtomicBC:
    """PeriodicBC(T) (bc_operators.jl:17-19)."""

    periodic = True

    def key(self):
        return ("periodic",)

    def __repr__(self):
        return "PeriodicBC()"

    def __mul__(self, u):
        # Q * u (bc_operators.jl:188-192, multi_dim_bc_operators.jl:210-252): lazy padding
        return _bc_times(self, u)


def _robin_atomic(l, r, dx, order=1):
    al, bl, gl = (float(v) for v in l)
    ar, br, gr = (float(v) for v in r)
    order = int(order)
    with np.errstate(divide="ignore", invalid="ignore"):
        if np.ndim(dx) == 0:
            # bc_operators.jl:26-39
            dx = float(dx)
            s = calculate_weights(1, 1.0, [float(k) for k in range(1, order + 2)])
            dl = _div(al * dx, bl) + s[0]
            dr = _div(ar * dx, br) - s[0]
            a_l = [_div(-s[k], dl) for k in range(1, order + 1)]
            a_r = [_div(s[k], dr) for k in range(order, 0, -1)]
            b_l = _div(gl, al + _div(bl * s[0], dx))
            b_r = _div(gr, ar - _div(br * s[0], dx))
        else:
            # bc_operators.jl:40-60
            dxv = np.asarray(dx, dtype=np.float64)
            ns = order + 1
            dx_l, dx_r = dxv[:ns], dxv[len(dxv) - ns:]
            s = calculate_weights(1, 1.0, [float(k) for k in range(1, ns + 1)])
            denom_l = al + _div(bl * s[0], float(dx_l[0]))
            denom_r = ar - _div(br * s[0], float(dx_r[ns - 1]))
            a_l = [_div(-bl * s[k], denom_l * float(dx_l[k])) for k in range(1, ns)]
            a_r = [_div(br * s[ns - 1 - m], denom_r * float(dx_r[m])) for m in range(order)]
            b_l = _div(gl, denom_l)
            b_r = _div(gr, denom_r)
    return AffineBC(a_l, b_l, a_r, b_r)


def _div(a, b):
    """IEEE a / b (Python raises on float division by zero; Julia returns +-Inf / NaN)."""
    return float(np.float64(a) / np.float64(b))


class MultiDimDirectionalBC:
    """An atomic BC extended over every line along axis D of an array of size s
    (multi_dim_bc_operators.jl:54-57, 97-101), or one atomic BC per line: `lines` is an array
    (numpy, dtype object) of AffineBC indexed like the array's other axes in Julia order
    (MultiDimBC{dim}(BC::Array), :93-95; line I gets BC[I], :34); then s is None."""

    def __init__(self, bc: AffineBC, D: int, s, lines=None):
        self.bc, self.D = bc, int(D)
        self.s = None if s is None else tuple(int(v) for v in s)
        self.lines = lines

    @property
    def ndims(self):
        return len(self.s) if self.s is not None else self.lines.ndim + 1

    def __add__(self, other):
        return compose(self, other)

    def __mul__(self, u):
        # Q * u (bc_operators.jl:188-192, multi_dim_bc_operators.jl:210-252): lazy padding
        return _bc_times(self, u)


class ComposedMultiDimBC:
    """One directional BC per axis (multi_dim_bc_operators.jl:59-61)."""

    def __init__(self, bcs):
        self.bcs = tuple(bcs)   # sorted by axis

    @property
    def s(self):
        return self.bcs[0].s

    @property
    def ndims(self):
        return self.bcs[0].ndims

    def __mul__(self, u):
        # Q * u (bc_operators.jl:188-192, multi_dim_bc_operators.jl:210-252): lazy padding
        return _bc_times(self, u)


def compose(*BCs):
    """multi_dim_bc_operators.jl:178-192."""
    N = BCs[0].ndims
    Ds = [b.D for b in BCs]
    if len(BCs) != N:
        raise ValueError("ArgumentError: There must be enough BCs to cover every dimension - "
                         "check that the number of MultiDimBCs == N")
    if len(set(Ds)) != len(Ds):
        raise ValueError("ArgumentError: There are multiple boundary conditions that extend "
                         "along the same dimension - make sure every dimension has a unique extension")
    return ComposedMultiDimBC(sorted(BCs, key=lambda b: b.D))


def decompose(Q: ComposedMultiDimBC):
    return Q.bcs


class BoundaryPadded:
    """Q * u: u with Q's ghosts, kept lazy like BoundaryPaddedVector / BoundaryPaddedArray /
    ComposedBoundaryPaddedArray (boundary_padded_arrays.jl:14-180).  L * (Q*u) and
    mul_(du, L, Q*u) run the same fused kernel as (L*Q) * u; materialize() builds the padded
    tensor (size + 2 along each padded axis, zero corners for a composed BC)."""

    __array_ufunc__ = None

    def __init__(self, Q, u):
        self.Q, self.u = Q, u

    def materialize(self):
        """The padded tensor (Array(Q*u)), assembled with torch from the ghost formulas."""
        import torch
        from .apply import _PeriodicPad, _bcs_of, julia_size
        u = self.u
        js = julia_size(u)
        nd = len(js)
        bcs, _ = _bcs_of(self.Q, nd, js)
        pshape = list(u.shape)
        for ax in bcs:
            pshape[nd - 1 - ax] += 2
        out = torch.zeros(pshape, dtype=u.dtype, device=u.device)
        inner = tuple(slice(1, s - 1) if (nd - 1 - k) in bcs else slice(None)
                      for k, s in enumerate(pshape))
        out[inner] = u
        for ax, bc in bcs.items():
            dim = nd - 1 - ax
            n = u.shape[dim]
            x = u.to(torch.float64).movedim(dim, -1)
            if isinstance(bc, _PeriodicPad):
                gl, gr = (x[..., n - 1], x[..., 0]) if bc.kind == 1 else (x[..., 0], x[..., n - 1])
            elif isinstance(bc, _LineBCs):
                al, bl, ar, br, per = (torch.as_tensor(t, device=u.device) for t in bc.tables())
                al, bl, ar, br = (t.to(torch.float64) for t in (al, bl, ar, br))
                xl = x.reshape(-1, n)   # C order over the other axes = Julia's column-major lines
                k = al.shape[1]
                gl = (xl[:, :k] * al).sum(-1) + bl
                k = ar.shape[1]
                gr = (xl[:, n - k:] * ar).sum(-1) + br
                gl = torch.where(per, xl[:, n - 1], gl).reshape(x.shape[:-1])
                gr = torch.where(per, xl[:, 0], gr).reshape(x.shape[:-1])
            else:
                al = torch.as_tensor(bc.a_l, dtype=torch.float64, device=u.device)
                ar = torch.as_tensor(bc.a_r, dtype=torch.float64, device=u.device)
                gl = (x[..., : len(bc.a_l)] * al).sum(-1) + bc.b_l
                gr = (x[..., n - len(bc.a_r):] * ar).sum(-1) + bc.b_r
            # ghost planes sit inside the other padded axes (zero corners)
            lo = [slice(1, s - 1) if ((nd - 1 - k) in bcs and k != dim) else slice(None)
                  for k, s in enumerate(pshape)]
            hi = list(lo)
            lo[dim], hi[dim] = 0, pshape[dim] - 1
            out[tuple(lo)] = gl.to(u.dtype)
            out[tuple(hi)] = gr.to(u.dtype)
        return out


class _LineBCs:
    """One atomic BC per line along an axis (lines in Julia column-major order): affine, or
    PeriodicBC, whose line wraps (slice_rmul pads line I with BC[I] * u[I, :],
    multi_dim_bc_operators.jl:29-39; bc_operators.jl:192)."""

    def __init__(self, lines):
        self.lines = lines

    def key(self):
        return ("lines", id(self.lines))

    def flat(self):
        return list(self.lines.flatten(order="F"))

    def tables(self):
        """(a_l, b_l, a_r, b_r, periodic) per line, taps right-aligned for a_r and
        zero-padded; periodic lines carry no taps."""
        bcs = self.flat()
        per = np.array([isinstance(b, PeriodicAtomicBC) for b in bcs])
        aff = [b for b in bcs if not isinstance(b, PeriodicAtomicBC)]
        kl = max([len(b.a_l) for b in aff], default=0)
        kr = max([len(b.a_r) for b in aff], default=0)
        al = np.zeros((len(bcs), kl))
        ar = np.zeros((len(bcs), kr))
        bl = np.zeros(len(bcs))
        br = np.zeros(len(bcs))
        for i, b in enumerate(bcs):
            if per[i]:
                continue
            al[i, : len(b.a_l)] = b.a_l
            if len(b.a_r):
                ar[i, kr - len(b.a_r):] = b.a_r
            bl[i], br[i] = b.b_l, b.b_r
        return al, bl, ar, br, per


def all_periodic_lines(lines) -> bool:
    """A BC array whose entries are all PeriodicBC (Julia eltype PeriodicBC{T}), which takes the
    N-D periodic padding (multi_dim_bc_operators.jl:221-228) instead of slice_rmul."""
    return all(isinstance(b, PeriodicAtomicBC) for b in np.asarray(lines, dtype=object).flat)


def _bc_times(Q, u):
    return BoundaryPadded(Q, u)


def is_bc(x) -> bool:
    return isinstance(x, (AffineBC, PeriodicAtomicBC, MultiDimDirectionalBC, ComposedMultiDimBC))


def _is_size(s):
    return isinstance(s, (tuple, list)) and all(isinstance(v, numbers.Integral) for v in s)


class _DimBC:
    """`Name(...)` and `Name[dim](...)` like Julia's `Name` / `Name{dim}`."""

    def __init__(self, fn, fn_dim):
        self._fn, self._fn_dim = fn, fn_dim

    def __call__(self, *a, **k):
        return self._fn(*a, **k)

    def __getitem__(self, dim):
        return lambda *a, **k: self._fn_dim(int(dim), *a, **k)


def _robin(l, r, dx, order=1, s=None):
    if s is None:
        return _robin_atomic(l, r, dx, order)
    # RobinBC(l, r, dxyz, order, s): one directional BC per axis (:141-149)
    return tuple(MultiDimDirectionalBC(_robin_atomic(l, r, dx[dim], order), dim + 1, s)
                 for dim in range(len(s)))


def _robin_dim(dim, l, r, dx, order, s):
    return MultiDimDirectionalBC(_robin_atomic(l, r, dx, order), dim, s)


def _dirichlet(al, ar, s=None):
    if s is None:   # bc_operators.jl:176-178
        return _robin_atomic((1.0, 0.0, al), (1.0, 0.0, ar), 1.0, 2)
    # multi_dim_bc_operators.jl:126-128: dxyz = fill(1, ndims)
    return _robin((1.0, 0.0, al), (1.0, 0.0, ar), [1.0] * len(s), 2, s)


def _dirichlet_dim(dim, al, ar, s):
    return _robin_dim(dim, (1.0, 0.0, al), (1.0, 0.0, ar), 1.0, 2, s)


def _dirichlet0(T=float, s=None):
    return _dirichlet(0.0, 0.0, s)


def _dirichlet0_dim(dim, T, s):
    return _dirichlet_dim(dim, 0.0, 0.0, s)


def _neumann(alpha, dx, order=1, s=None):
    return _robin((0.0, 1.0, alpha[0]), (0.0, 1.0, alpha[1]), dx, order, s)


def _neumann_dim(dim, alpha, dx, order, s):
    return _robin_dim(dim, (0.0, 1.0, alpha[0]), (0.0, 1.0, alpha[1]), dx, order, s)


def _neumann0(*args):
    # Neumann0BC(dx, order=1)  |  Neumann0BC(T, dxyz, order, s)
    if len(args) == 4:
        _, dxyz, order, s = args
        return _neumann((0.0, 0.0), dxyz, order, s)
    dx = args[0]
    order = args[1] if len(args) > 1 else 1
    return _neumann((0.0, 0.0), dx, order)


def _neumann0_dim(dim, T, dx, order, s):
    return _neumann_dim(dim, (0.0, 0.0), dx, order, s)


def _multidim(bc, s):
    if not _is_size(s):
        raise TypeError("MultiDimBC(bc, s) needs the size tuple s")
    return tuple(MultiDimDirectionalBC(bc, dim + 1, s) for dim in range(len(s)))


def _multidim_dim(dim, bc, s=None):
    if s is None:
        # MultiDimBC{dim}(BC::Array{<:AtomicBC}) (multi_dim_bc_operators.jl:93-95)
        lines = np.asarray(bc, dtype=object)
        if lines.ndim < 1 or lines.ndim > 2:
            raise ValueError("the BC array must have 1 or 2 dimensions (2-D or 3-D arrays)")
        for b in lines.flat:
            if not isinstance(b, (AffineBC, PeriodicAtomicBC)):
                raise TypeError("per-line boundary conditions must be atomic (Robin / Dirichlet / "
                                "Neumann / GeneralBC / PeriodicBC)")
        return MultiDimDirectionalBC(lines.flat[0], dim, None, lines=lines)
    return MultiDimDirectionalBC(bc, dim, s)


RobinBC = _DimBC(_robin, _robin_dim)
DirichletBC = _DimBC(_dirichlet, _dirichlet_dim)
Dirichlet0BC = _DimBC(_dirichlet0, _dirichlet0_dim)
NeumannBC = _DimBC(_neumann, _neumann_dim)
Neumann0BC = _DimBC(_neumann0, _neumann0_dim)
MultiDimBC = _DimBC(_multidim, _multidim_dim)


def _periodic(T=float, s=None):
    if s is None:
        return PeriodicAtomicBC()
    # PeriodicBC(T, s) = MultiDimBC(PeriodicBC(T), s)  (multi_dim_bc_operators.jl:114)
    return _multidim(PeriodicAtomicBC(), s)


def _periodic_dim(dim, T, s):
    # PeriodicBC{dim}(T, s) = MultiDimBC{dim}(PeriodicBC(T), s)  (:113)
    return MultiDimDirectionalBC(PeriodicAtomicBC(), dim, s)


PeriodicBC = _DimBC(_periodic, _periodic_dim)


def _ipow(x, i):
    r = x
    for _ in range(1, i):
        r = r * x
    return r


def _general_side(alpha, dx, order, right):
    """One end of GeneralBC (bc_operators.jl:85-170): alpha[0] + alpha[1] u + sum_i alpha[i+1]
    u^(i) = 0, the i-th derivative taken on the points 1..order+i (at point 1 on the left, at
    point order+i on the right), rows zero-padded at the end as in the reference."""
    alpha = [float(v) for v in alpha]
    n = len(alpha)
    ncol = order + n - 2
    vec = np.ndim(dx) != 0
    dxv = np.asarray(dx, dtype=np.float64) if vec else float(dx)
    S = []
    for i in range(1, n - 1):
        m = order + i
        w = calculate_weights(i, float(m) if right else 1.0, [float(k) for k in range(1, m + 1)])
        row = []
        for j in range(ncol):
            v = float(w[j]) if j < m else 0.0
            if vec:
                h = _ipow(float(dxv[len(dxv) - 1 - j]) if right else float(dxv[j]), i)
            else:
                h = _ipow(dxv, i)
            row.append(_div(v, h))
        S.append(row)
    denom = alpha[1]
    for i in range(n - 2):
        denom = denom + alpha[i + 2] * S[i][ncol - 1 if right else 0]
    nt = ncol - 1
    t = []
    for j in range(nt):
        col = (ncol - 2 - j) if right else (j + 1)
        acc = 0.0
        for i in range(n - 2):
            acc = acc + alpha[i + 2] * S[i][col]
        t.append(_div(-acc, denom))
    a = list(reversed(t)) if right else t
    return a, _div(-alpha[0], denom)


def GeneralBC(alpha_l, alpha_r, dx, order=1):
    """GeneralBC(αl, αr, dx, order) (bc_operators.jl:85-170): an affine BC whose ghost solves
    αl[1] + αl[2] u + αl[3] u' + αl[4] u'' + ... = 0 at the left end (likewise at the right)."""
    if len(alpha_l) < 2 or len(alpha_r) < 2:
        raise ValueError("GeneralBC needs at least two coefficients per side")
    if np.ndim(dx) != 0:
        need = max(order + len(alpha_l) - 2, order + len(alpha_r) - 2)
        if len(dx) < need:
            raise ValueError("GeneralBC: dx vector too short")
    with np.errstate(divide="ignore", invalid="ignore"):
        a_l, b_l = _general_side(alpha_l, dx, order, False)
        a_r, b_r = _general_side(alpha_r, dx, order, True)
    return AffineBC(a_l, b_l, a_r, b_r)
