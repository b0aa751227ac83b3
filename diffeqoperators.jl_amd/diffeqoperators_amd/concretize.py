"""Concretization and direct solves, on the device.

Mirrors concretization.jl (`Array`, `sparse` of operators and of L*Q) and
ghost_derivative_operator.jl:39-58 (`A \\ u`).  The matrices are exported by libdop from the
plan the operator applies (dop_plan_csr_rows / dop_plan_csr_fill), so `A*u == Array(A)*u`
holds by construction; the solve factors the band of L*Q_L on the device
(dop_solver_create) and solves there (dop_solver_solve).

Sizes follow the reference: `sparse(L, N)` / `Array(L, N)` take the line length of a
DerivativeOperator (N = L.len), `sparse(L, Mshape)` the Julia size of the padded array M it is
applied to (the Kron form, concretization.jl:425-444); for L*Q, `s` is the unpadded size of u.
Matrices are FP64 torch tensors on the GPU: `sparse` returns a CSR tensor, `Array` a dense
one; for L*Q both return the pair (A_l, A_b).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from .apply import _ghost_terms, _padded_plan, ghost_plans
from .composite import DiffEqOperatorCombination
from .operators import DerivativeOperator

__all__ = ["sparse", "Array", "ldiv", "Solver"]


def _fake(shape, dtype=torch.float64, device=None):
    """A tensor of the given shape that owns one element (only its shape, dtype and device
    are read by the plan builders)."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    return torch.empty(1, dtype=dtype, device=device).as_strided(tuple(shape), (0,) * len(shape))


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def plan_csr(plan, device):
    """(row_ptr, col, val, b, kl, ku) of one plan: du = A * vec(src) + b."""
    L = _abi.lib()
    n = plan.out_elems
    rp = torch.empty(n + 1, dtype=torch.int64, device=device)
    nnz, kl, ku = C.c_int64(), C.c_int64(), C.c_int64()
    s = C.c_void_p(_stream(device))
    _abi.check(L.dop_plan_csr_rows(plan.handle, C.c_void_p(rp.data_ptr()), C.byref(nnz),
                                   C.byref(kl), C.byref(ku), s), "dop_plan_csr_rows")
    col = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=device)[: nnz.value]
    val = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=device)[: nnz.value]
    b = torch.empty(n, dtype=torch.float64, device=device)
    _abi.check(L.dop_plan_csr_fill(plan.handle, C.c_void_p(rp.data_ptr()),
                                   C.c_void_p(col.data_ptr()), C.c_void_p(val.data_ptr()),
                                   C.c_void_p(b.data_ptr()), s), "dop_plan_csr_fill")
    return rp, col, val, b, kl.value, ku.value


def _csr_tensor(rp, col, val, n_rows, n_cols):
    return torch.sparse_csr_tensor(rp, col, val, size=(n_rows, n_cols), dtype=torch.float64,
                                   device=rp.device)


def _sum_csr(parts, n_rows, n_cols):
    if len(parts) == 1:
        rp, col, val = parts[0]
        return _csr_tensor(rp, col, val, n_rows, n_cols)
    coo = None
    for rp, col, val in parts:
        t = _csr_tensor(rp, col, val, n_rows, n_cols).to_sparse_coo()
        coo = t if coo is None else coo + t
    return coo.coalesce().to_sparse_csr()


def _ghost_size(A, s):
    L0 = _ghost_terms(A)[0][0]
    if s is None:
        return (L0.len,)
    if isinstance(s, int):
        return (s,)
    return tuple(int(v) for v in s)


def _concretize(A, s, device):
    """(A_l CSR, A_b or None)."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    terms = _ghost_terms(A)
    if terms is not None:
        jsize = _ghost_size(A, s)
        u = _fake(tuple(reversed(jsize)), device=device)
        plans = ghost_plans(A, terms, u)
        parts, bsum = [], None
        for p in plans:
            rp, col, val, b, _, _ = plan_csr(p, device)
            parts.append((rp, col, val))
            bsum = b if bsum is None else bsum + b
        n = plans[0].out_elems
        return _sum_csr(parts, n, n), bsum
    if isinstance(A, DerivativeOperator):
        ops = [A]
    elif isinstance(A, DiffEqOperatorCombination) and all(
            isinstance(op, DerivativeOperator) for op in A.ops):
        ops = list(A.ops)
        if len({op.N for op in ops}) != 1:
            # concretization.jl:918-924 sums per-operator Kron matrices whose row sets differ
            raise ValueError("DimensionMismatch: operators along different axes concretize to "
                             "matrices of different sizes")
    else:
        raise TypeError(f"cannot concretize {type(A)}")
    N = ops[0].N
    if s is None or isinstance(s, int):
        n = ops[0].len if s is None else int(s)
        if n != ops[0].len:
            raise ValueError(f"the operator is defined on lines of length {ops[0].len}, not {n}")
        if N != 1:
            raise ValueError("give the padded array size Mshape for an operator along axis > 1")
        msz = (n + 2,)
    else:
        msz = tuple(int(v) for v in s)
    if N > len(msz):
        raise AssertionError(f"operator axis {N} > ndims {len(msz)}")
    osz = list(msz)
    osz[N - 1] -= 2
    du = _fake(tuple(reversed(osz)), device=device)
    M = _fake(tuple(reversed(msz)), device=device)
    plan = _padded_plan(A, ops, du, M)
    rp, col, val, _, _, _ = plan_csr(plan, device)
    return _csr_tensor(rp, col, val, plan.out_elems, plan.in_elems), None


def sparse(A, s=None, *, device=None):
    """sparse(A, N | s) (concretization.jl): a CSR tensor for a DerivativeOperator (or a sum
    along one axis), the pair (A_l, A_b) for L*Q (and sums of them):
    `A_l @ vec(u) + A_b == vec(A*u)`, vec = column-major (Julia) order."""
    A_l, A_b = _concretize(A, s, device)
    return A_l if A_b is None else (A_l, A_b)


def Array(A, s=None, *, device=None):
    """Array(A, N | s) (concretization.jl): the dense form of `sparse(A, s)`."""
    A_l, A_b = _concretize(A, s, device)
    return A_l.to_dense() if A_b is None else (A_l.to_dense(), A_b)


class Solver:
    """The device factorization of a square L*Q (or a sum of L*Q with one BC per axis):
    `solve(rhs)` returns A \\ rhs = A_l \\ (rhs - A_b) (ghost_derivative_operator.jl:39-58).
    rhs: float64 tensor of the array's shape, or a stack of them (leading batch axis)."""

    def __init__(self, A, shape, device=None):
        terms = _ghost_terms(A)
        if terms is None:
            raise TypeError("A \\ u needs L*Q (a GhostDerivativeOperator) or a sum of them")
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.shape = tuple(shape)
        plans = ghost_plans(A, terms, _fake(self.shape, device=device))
        if len(plans) != 1:
            raise ValueError("A \\ u needs the operators' BCs to agree on every axis")
        self._plan = plans[0]
        self.device = device
        h = C.c_void_p()
        with torch.cuda.device(device):
            _abi.check(_abi.lib().dop_solver_create(self._plan.handle, C.c_void_p(_stream(device)),
                                                    C.byref(h)), "dop_solver_create")
        self.handle = h
        n, kl, ku = C.c_int64(), C.c_int64(), C.c_int64()
        _abi.check(_abi.lib().dop_solver_band(h, C.byref(n), C.byref(kl), C.byref(ku)))
        self.n, self.kl, self.ku = n.value, kl.value, ku.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_solver_destroy(h)
            except Exception:
                pass
            self.handle = None

    def solve(self, rhs, out=None):
        if not isinstance(rhs, torch.Tensor) or rhs.device.type != "cuda":
            raise TypeError("rhs must be a torch tensor on the GPU (no CPU path)")
        if rhs.dtype != torch.float64:
            raise TypeError("A \\ u is computed in float64: pass a float64 right-hand side")
        if not rhs.is_contiguous():
            raise ValueError("rhs must be contiguous")
        batched = tuple(rhs.shape) != self.shape
        if batched and tuple(rhs.shape[1:]) != self.shape:
            raise AssertionError(f"rhs of shape {tuple(rhs.shape)} does not match {self.shape}")
        nrhs = rhs.shape[0] if batched else 1
        x = torch.empty_like(rhs) if out is None else out
        if x.shape != rhs.shape or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError("out must be a contiguous float64 tensor shaped like rhs")
        if x.data_ptr() == rhs.data_ptr():
            raise ValueError("out must not alias rhs")
        _abi.check(_abi.lib().dop_solver_solve(self.handle, C.c_void_p(x.data_ptr()),
                                               C.c_void_p(rhs.data_ptr()), int(nrhs),
                                               C.c_void_p(_stream(rhs.device))), "dop_solver_solve")
        return x


def ldiv(A, u):
    """A \\ u (ghost_derivative_operator.jl:39-58).  The factorization is cached on A per
    array shape."""
    if not isinstance(u, torch.Tensor) or u.device.type != "cuda":
        raise TypeError("u must be a torch tensor on the GPU (no CPU path)")
    cache = getattr(A, "_solvers", None)
    if cache is None:
        cache = {}
        try:
            A._solvers = cache
        except AttributeError:
            pass
    key = (tuple(u.shape), u.device)
    if key not in cache:
        cache[key] = Solver(A, u.shape, u.device)
    return cache[key].solve(u)
