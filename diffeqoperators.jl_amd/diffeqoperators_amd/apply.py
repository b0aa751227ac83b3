"""Lowering of operators/BCs to libdop plans and their application on device tensors.

Replaces the reference's evaluation entry points (derivative_operator_functions.jl:18-69,
150-163, 203-802; ghost_derivative_operator.jl:15-37; composite_operators.jl:64-84):
every call becomes one fused dop_apply on the current torch stream.  Tensors must be
contiguous device tensors (float64, float32 or bfloat16); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

from . import _abi
from .bcs import (AffineBC, ComposedMultiDimBC, MultiDimDirectionalBC, PeriodicAtomicBC, _LineBCs,
                  all_periodic_lines)
from .composite import DiffEqOperatorCombination, GhostDerivativeOperator
from .operators import DerivativeOperator

_STORAGE = {torch.float64: _abi.F64, torch.float32: _abi.F32, torch.bfloat16: _abi.BF16}


def storage_of(t: torch.Tensor) -> int:
    try:
        return _STORAGE[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}: use float64, float32 or bfloat16") from None


def _check_device_tensor(t, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU (no CPU path)")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must live on the GPU (got {t.device}); this library has no CPU path")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def julia_size(t: torch.Tensor):
    """Column-major size of a tensor (Julia's size(u))."""
    return tuple(int(v) for v in reversed(t.shape))


class Plan:
    """An immutable libdop plan (dop_plan_create)."""

    def __init__(self, desc: _abi.DopDesc, keep):
        self._keep = keep
        h = C.c_void_p()
        _abi.check(_abi.lib().dop_plan_create(C.byref(desc), C.byref(h)), "dop_plan_create")
        self.handle = h
        self.storage = desc.storage
        self.in_elems = _abi.lib().dop_plan_input_elems(h)
        self.out_elems = _abi.lib().dop_plan_output_elems(h)
        self.desc = desc

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _abi.lib().dop_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def apply(self, du: torch.Tensor, u: torch.Tensor, accumulate: bool = False, stream=None):
        if du.numel() != self.out_elems or u.numel() != self.in_elems:
            raise AssertionError("array sizes do not match the plan")
        if storage_of(du) != self.storage or storage_of(u) != self.storage:
            raise TypeError("tensor dtype does not match the plan's storage type")
        if du.data_ptr() == u.data_ptr():
            raise ValueError("du must not alias u")
        if stream is None:
            stream = torch.cuda.current_stream(u.device).cuda_stream
        _abi.check(_abi.lib().dop_apply(self.handle, C.c_void_p(du.data_ptr()),
                                        C.c_void_p(u.data_ptr()), int(bool(accumulate)),
                                        C.c_void_p(stream)), "dop_apply")

    def update(self, desc: _abi.DopDesc, keep, stream=None):
        """New BC values / operator tables of the same structure, in place
        (dop_plan_update): the plan's device memory is reused.  Raises DopError (EINVAL) when
        the structure differs."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _abi.check(_abi.lib().dop_plan_update(self.handle, C.byref(desc), C.c_void_p(stream)),
                   "dop_plan_update")
        self._keep = keep
        self.desc = desc

    @property
    def variant(self):
        """(candidate launched, tuned?) -- dop_plan_variant."""
        t = C.c_int()
        v = _abi.lib().dop_plan_variant(self.handle, C.byref(t))
        return v, bool(t.value)

    def autotune(self, stream=None):
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        choice = C.c_int()
        _abi.check(_abi.lib().dop_plan_autotune(self.handle, C.c_void_p(stream), C.byref(choice)),
                   "dop_plan_autotune")
        return choice.value


def make_desc(terms, jsize, storage, *, input_kind, semantics, bcs=None, padded=(0, 0, 0),
              transverse_view=0, accum=_abi.ACC_AUTO):
    """terms: list of DerivativeOperator; bcs: {axis0: AffineBC}."""
    ndim = len(jsize)
    if ndim < 1 or ndim > 3:
        raise ValueError("arrays of 1 to 3 dimensions are supported")
    ops = (_abi.DopOperator * len(terms))()
    keep = [ops]
    for t, L in enumerate(terms):
        if not L.N <= ndim:
            raise AssertionError(f"operator axis {L.N} > ndims {ndim}")
        ops[t] = L.c_struct(L.N - 1)
        keep.append(L)
    d = _abi.DopDesc()
    d.ndim = ndim
    d.storage = storage
    d.accum = accum
    d.input = input_kind
    d.semantics = semantics
    d.n_ops = len(terms)
    for k in range(3):
        d.size[k] = jsize[k] if k < ndim else 1
        d.padded[k] = int(padded[k]) if k < ndim else 0
    d.transverse_view = int(transverse_view)
    d.ops = ops
    for ax, bc in (bcs or {}).items():
        q = d.bc[ax]
        if isinstance(bc, _PeriodicPad):
            q.present, q.kind, q.n_l, q.n_r = 1, bc.kind, 0, 0
            continue
        if isinstance(bc, _LineBCs):
            # one BC per line: the axis entry only marks the axis as padded
            q.present, q.kind, q.n_l, q.n_r = 1, _abi.BC_AFFINE, 0, 0
            flat = bc.flat()
            arr = (_abi.DopAffineBC * len(flat))()
            for i, b in enumerate(flat):
                if isinstance(b, PeriodicAtomicBC):
                    arr[i].present, arr[i].kind = 1, _abi.BC_PERIODIC
                    continue
                if len(b.a_l) > _abi.MAX_BC_TAPS or len(b.a_r) > _abi.MAX_BC_TAPS:
                    raise ValueError("boundary condition order too high")
                e = arr[i]
                e.present, e.kind, e.n_l, e.n_r = 1, _abi.BC_AFFINE, len(b.a_l), len(b.a_r)
                for k, v in enumerate(b.a_l):
                    e.a_l[k] = float(v)
                for k, v in enumerate(b.a_r):
                    e.a_r[k] = float(v)
                e.b_l, e.b_r = float(b.b_l), float(b.b_r)
            d.line_bc[ax] = C.cast(arr, C.POINTER(_abi.DopAffineBC))
            keep.append(arr)
            continue
        if len(bc.a_l) > _abi.MAX_BC_TAPS or len(bc.a_r) > _abi.MAX_BC_TAPS:
            raise ValueError("boundary condition order too high")
        q.present = 1
        q.n_l, q.n_r = len(bc.a_l), len(bc.a_r)
        for k, v in enumerate(bc.a_l):
            q.a_l[k] = float(v)
        for k, v in enumerate(bc.a_r):
            q.a_r[k] = float(v)
        q.b_l, q.b_r = bc.b_l, bc.b_r
    return d, keep


# ------------------------------------------------------------------------------------------
# Q*u: which BC pads which axis, and which reference kernels apply

class _PeriodicPad:
    """Periodic padding of one axis: kind BC_PERIODIC (line wrap, bc_operators.jl:192) or
    BC_PERIODIC_ND (first/last planes, multi_dim_bc_operators.jl:221-252)."""

    def __init__(self, kind):
        self.kind = kind

    def key(self):
        return ("periodic", self.kind)


def _bcs_of(Q, ndim, jsize):
    """{axis0: AffineBC | _PeriodicPad} of Q applied to an array of Julia size jsize, and the
    semantics."""
    if isinstance(Q, (AffineBC, PeriodicAtomicBC)):
        # 1-D: Q*u is a BoundaryPaddedVector (bc_operators.jl:188-192);
        # N-D: BC*u = MultiDimBC{1}(BC, size(u))*u pads axis 1 (multi_dim_bc_operators.jl:210),
        # which for PeriodicBC takes the N-D periodic method (:221-228)
        if isinstance(Q, PeriodicAtomicBC):
            Q = _PeriodicPad(_abi.BC_PERIODIC if ndim == 1 else _abi.BC_PERIODIC_ND)
        return {0: Q}, (_abi.SEM_BPV if ndim == 1 else _abi.SEM_LINES)
    if ndim == 1:
        raise AssertionError("multi-dimensional BCs need arrays of 2 or more dimensions")
    if isinstance(Q, MultiDimDirectionalBC):
        _check_dir_size(Q, jsize, "Size of the BCs array in the MultiDimBC is incorrect")
        if Q.lines is not None:
            if all_periodic_lines(Q.lines):
                return {Q.D - 1: _PeriodicPad(_abi.BC_PERIODIC_ND)}, _abi.SEM_LINES
            return {Q.D - 1: _LineBCs(Q.lines)}, _abi.SEM_LINES
        bc = Q.bc
        if isinstance(bc, PeriodicAtomicBC):
            bc = _PeriodicPad(_abi.BC_PERIODIC_ND)
        return {Q.D - 1: bc}, _abi.SEM_LINES
    if isinstance(Q, ComposedMultiDimBC):
        # all atoms periodic: the N-D periodic method (:244-252); otherwise every line is
        # padded by its own atom (slice_rmul), i.e. the line wrap for periodic atoms
        all_periodic = all(all_periodic_lines(b.lines) if b.lines is not None
                           else isinstance(b.bc, PeriodicAtomicBC) for b in Q.bcs)
        out = {}
        for b in Q.bcs:
            _check_dir_size(b, jsize, f"Size of the BCs array for dimension {b.D} is incorrect")
            if b.lines is not None and not all_periodic:
                out[b.D - 1] = _LineBCs(b.lines)
                continue
            bc = b.bc
            if isinstance(bc, PeriodicAtomicBC):
                bc = _PeriodicPad(_abi.BC_PERIODIC_ND if all_periodic else _abi.BC_PERIODIC)
            out[b.D - 1] = bc
        return out, _abi.SEM_LINES
    raise TypeError(f"not a boundary condition: {type(Q)}")


def _check_dir_size(Q, jsize, msg):
    """The BC array of a MultiDimDirectionalBC must match the array's other axes
    (multi_dim_bc_operators.jl:178-192 / perpindex)."""
    if Q.lines is not None:
        other = tuple(v for k, v in enumerate(jsize) if k != Q.D - 1)
        if len(jsize) != Q.lines.ndim + 1 or tuple(Q.lines.shape) != other:
            raise AssertionError(msg)
    elif tuple(Q.s) != tuple(jsize):
        raise AssertionError(msg)


def _ghost_groups(terms, ndim, jsize):
    """Split (L, Q) terms into groups whose per-axis BCs agree (one plan per group)."""
    groups = []   # [ops, bcs{axis: bc}, keys{axis: key}, sem]
    for L, Q in terms:
        bcs, sem = _bcs_of(Q, ndim, jsize)
        ax = L.N - 1
        if ax not in bcs:
            raise AssertionError("the operator's axis is not padded by its boundary condition")
        key = bcs[ax].key()
        for g in groups:
            if g[3] == sem and g[2].get(ax, key) == key:
                g[0].append(L)
                g[1][ax] = bcs[ax]
                g[2][ax] = key
                break
        else:
            groups.append([[L], {ax: bcs[ax]}, {ax: key}, sem])
    return groups


def _plan_cache(obj):
    c = getattr(obj, "_plans", None)
    if c is None:
        c = {}
        try:
            obj._plans = c
        except AttributeError:
            pass
    return c


def _op_structure(L):
    return (L.N, int(L.wind), L.derivative_order, L.approximation_order, L.len,
            L.stencil_length, L.boundary_stencil_length, L.boundary_point_count, L.offside,
            int(L.uniform_grid), L.coeff_kind)


def _bc_structure(bc):
    if isinstance(bc, _PeriodicPad):
        return bc.key()
    if isinstance(bc, _LineBCs):
        return ("lines", hash(tuple(("p",) if isinstance(b, PeriodicAtomicBC) else
                                    (len(b.a_l), len(b.a_r)) for b in bc.flat())))
    return ("affine", len(bc.a_l), len(bc.a_r))


class _Bound:
    """The plans of one owner (an L*Q object) and the descriptors of its values."""

    def __init__(self, skey, plans, descs):
        self.skey, self.plans, self.descs = skey, plans, descs


# Plans shared by owners of the same structure (same operators' kinds and sizes, same BC kinds
# and tap counts): a method-of-lines RHS that builds a new BC at every call (heat_equation.jl:
# 38-43, KdV.jl:26-30) reuses one plan and only replaces its values (dop_plan_update) instead
# of creating a plan per call.  Bounded: the least recently used structures are dropped.
_SHARED: "collections.OrderedDict" = None
_SHARED_MAX = 64


def _shared():
    global _SHARED
    if _SHARED is None:
        import collections
        _SHARED = collections.OrderedDict()
    return _SHARED


def ghost_plans(owner, terms, u: torch.Tensor, accum=_abi.ACC_AUTO, shared=False):
    """Plans evaluating sum_t L_t*Q_t*u, cached on owner.  shared=True (the implicit `A*u` /
    `mul_` paths, which use the plans at once): plans of the same structure are shared between
    owners and rebound to an owner's values on use.  shared=False (plans handed to the caller,
    who may keep them, e.g. in a batch or a server): the owner's own plans."""
    jsize = julia_size(u)
    storage = storage_of(u)
    key = ("ghost", jsize, storage, accum, u.device.index)
    cache = _plan_cache(owner)
    if not shared:
        own = cache.get(key + ("own",))
        if own is None:
            with torch.cuda.device(u.device):
                own = []
                for ops, bcs, _, sem in _ghost_groups(terms, len(jsize), jsize):
                    d, keep = make_desc(ops, jsize, storage, input_kind=_abi.INPUT_GHOST,
                                        semantics=sem, bcs=bcs, accum=accum)
                    own.append(Plan(d, keep + [bcs]))
            cache[key + ("own",)] = own
        return own
    bound = cache.get(key)
    if bound is None:
        descs, groups = [], _ghost_groups(terms, len(jsize), jsize)
        for ops, bcs, _, sem in groups:
            d, keep = make_desc(ops, jsize, storage, input_kind=_abi.INPUT_GHOST,
                                semantics=sem, bcs=bcs, accum=accum)
            descs.append((d, keep + [bcs]))
        # plans are created with the kernel selection the environment asks for
        env = (os.environ.get("DOP_FORCE_GENERIC"), os.environ.get("DOP_TILE_VARIANT"))
        skey = key + env + (tuple((sem, tuple(_op_structure(L) for L in ops),
                             tuple(sorted((ax, _bc_structure(bc)) for ax, bc in bcs.items())))
                            for ops, bcs, _, sem in groups),)
        bound = cache[key] = _Bound(skey, None, descs)
    shared = _shared()
    plans = shared.get(bound.skey)
    with torch.cuda.device(u.device):
        if plans is None:
            plans = [Plan(d, keep) for d, keep in bound.descs]
            for p in plans:
                p._owner = bound
            shared[bound.skey] = plans
            while len(shared) > _SHARED_MAX:
                shared.popitem(last=False)
        else:
            shared.move_to_end(bound.skey)
            for i, p in enumerate(plans):
                if p._owner is bound:
                    continue
                d, keep = bound.descs[i]
                try:
                    p.update(d, keep)
                except _abi.DopError:
                    # same kinds and sizes but a different row structure (e.g. coefficients
                    # that are no longer uniform): a plan of its own
                    p = plans[i] = Plan(d, keep)
                p._owner = bound
    bound.plans = plans
    return plans


def _ghost_terms(A):
    if isinstance(A, GhostDerivativeOperator):
        return [(A.L, A.Q)]
    if isinstance(A, DiffEqOperatorCombination) and all(
            isinstance(op, GhostDerivativeOperator) for op in A.ops):
        return [(op.L, op.Q) for op in A.ops]
    return None


def _apply_plans(plans, du, u, accumulate):
    for i, p in enumerate(plans):
        p.apply(du, u, accumulate=(accumulate or i > 0))


def ghost_times(G, u):
    _check_device_tensor(u, "u")
    out = torch.empty_like(u)
    _apply_plans(ghost_plans(G, _ghost_terms(G), u, shared=True), out, u, False)
    return out


# ------------------------------------------------------------------------------------------
# operators on padded arrays

def _padded_plan(owner, ops, du, M, transverse_view=0, accum=_abi.ACC_AUTO):
    osz, msz = julia_size(du), julia_size(M)
    if len(osz) != len(msz):
        raise AssertionError("du and M must have the same number of dimensions")
    padded = []
    for k in range(len(osz)):
        if msz[k] == osz[k]:
            padded.append(0)
        elif msz[k] == osz[k] + 2:
            padded.append(1)
        else:
            raise AssertionError("M must equal du in size or be padded by 2 along each axis")
    for L in ops:
        if not L.N <= len(osz) or not padded[L.N - 1]:
            raise AssertionError("differentiated dimension must be padded")
    storage = storage_of(M)
    key = ("padded", tuple(id(L) for L in ops), osz, msz, storage, transverse_view, accum)
    cache = _plan_cache(owner)
    if key not in cache:
        with torch.cuda.device(M.device):
            d, keep = make_desc(list(ops), osz, storage, input_kind=_abi.INPUT_PADDED,
                                semantics=_abi.SEM_LINES, padded=padded + [0] * (3 - len(padded)),
                                transverse_view=transverse_view, accum=accum)
            cache[key] = Plan(d, keep)
    return cache[key]


def derivative_times_array(L: DerivativeOperator, M):
    """L * M (derivative_operator_functions.jl:150-163): M is padded along L's axis."""
    _check_device_tensor(M, "M")
    shape = list(M.shape)
    shape[len(shape) - L.N] -= 2
    out = torch.empty(shape, dtype=M.dtype, device=M.device)
    _padded_plan(L, [L], out, M).apply(out, M)
    return out


def combination_times(A: DiffEqOperatorCombination, x):
    """(L1 + L2 + ...) * x = sum(op * x) (composite_operators.jl:64-65)."""
    _check_device_tensor(x, "x")
    terms = _ghost_terms(A)
    if terms is not None:
        out = torch.empty_like(x)
        _apply_plans(ghost_plans(A, terms, x, shared=True), out, x, False)
        return out
    if all(isinstance(op, DerivativeOperator) for op in A.ops):
        shapes = set()
        for L in A.ops:
            s = list(x.shape)
            s[len(s) - L.N] -= 2
            shapes.add(tuple(s))
        if len(shapes) != 1:
            raise ValueError("DimensionMismatch: the operators' results have different sizes")
        out = torch.empty(shapes.pop(), dtype=x.dtype, device=x.device)
        _padded_plan(A, A.ops, out, x).apply(out, x)
        return out
    parts = [op * x for op in A.ops]
    out = parts[0]
    for p in parts[1:]:
        out = out + p
    return out


def mul_(du, A, u, overwrite=True, accum=_abi.ACC_AUTO):
    """mul!(du, A, u): writes A*u into du (adds into du when overwrite=False)."""
    from .bcs import BoundaryPadded
    if isinstance(u, BoundaryPadded):
        # mul!(du, L, Q*u) = mul!(du, L*Q, u)
        return mul_(du, A * u.Q, u.u, overwrite=overwrite, accum=accum)
    _check_device_tensor(du, "du")
    _check_device_tensor(u, "u")
    accumulate = not overwrite
    terms = _ghost_terms(A)
    if terms is not None:
        if tuple(du.shape) != tuple(u.shape):
            raise AssertionError("du must have the unpadded size of u")
        _apply_plans(ghost_plans(A, terms, u, accum, shared=True), du, u, accumulate)
        return du
    if isinstance(A, DerivativeOperator):
        # derivative_operator_functions.jl:18-69
        _padded_plan(A, [A], du, u, accum=accum).apply(du, u, accumulate=accumulate)
        return du
    if isinstance(A, DiffEqOperatorCombination) and all(
            isinstance(op, DerivativeOperator) for op in A.ops):
        nd = du.dim()
        if nd == 1:
            tv = 0   # composite_operators.jl:76-84: plain per-op mul! and sum
        elif nd == 2:
            axes = {op.N for op in A.ops}
            tv = 1 if len(axes) > 1 else 0   # derivative_operator_functions.jl:414-461
        else:
            tv = 1                           # derivative_operator_functions.jl:768-801
        _padded_plan(A, A.ops, du, u, transverse_view=tv, accum=accum).apply(
            du, u, accumulate=accumulate)
        return du
    raise TypeError(f"mul_ does not support {type(A)}")


def plans_for(A, u, accum=_abi.ACC_AUTO):
    """The libdop plans that `A * u` runs (for benchmarking and batching)."""
    terms = _ghost_terms(A)
    if terms is None:
        raise TypeError("plans_for needs L*Q or a sum of L*Q")
    return ghost_plans(A, terms, u, accum)


def plan_for_shape(A, shape, dtype=torch.float32, device="cuda", accum=_abi.ACC_AUTO):
    """The single fused plan of L*Q (or a sum of them) for an array of the given tensor
    shape and dtype, without needing the array itself."""
    u = torch.empty(1, dtype=dtype, device=device).as_strided(tuple(shape), (0,) * len(shape))
    terms = _ghost_terms(A)
    if terms is None:
        raise TypeError("plan_for_shape needs L*Q or a sum of L*Q")
    plans = ghost_plans(A, terms, u, accum)
    if len(plans) != 1:
        raise ValueError("the operator needs one BC per axis to fuse into one plan")
    return plans[0]
