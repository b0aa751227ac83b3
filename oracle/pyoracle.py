"""ctypes binding of the FP64 CPU restatement (oracle/dop_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.  The product package never
imports this module.

Arrays follow the reference's column-major order: a Julia array of size (n1, n2, n3)
is a C-contiguous numpy array of shape (n3, n2, n1).

Beyond the binding, restated here in numpy and pinned by the tests named with each:
tsit5 (tests/test_oracle_tsit5.py: order conditions, matrix exponential), gradient /
divergence / curl (tests/test_oracle_vector_calculus.py: the reference tests' analytic
fields, curl of a gradient, a planted sign mistake), nonlinear_diffusion
(tests/test_oracle_nonlinear.py: Fast_Diffusion.jl's closed form, p = 1, a planted swap).
Parity unpinned: OrdinaryDiffEq's own Tsit5 step sequence (an absent dependency with no pinned
version; only the published method is restated) and the vector-calculus transverse view 1:n on
non-separable 2-D / 3-D fields (restated as read; the reference's tests use separable fields).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("DOP_ORACLE_LIBRARY") or os.path.join(_HERE, "_build", "libdop_oracle.so")

COEFF_VECTOR, COEFF_NUM_POS, COEFF_NUM_NEG = 0, 1, 2

_dp = C.POINTER(C.c_double)


class OrcOp(C.Structure):
    _fields_ = [
        ("axis", C.c_int), ("wind", C.c_int), ("d", C.c_int), ("a", C.c_int),
        ("len", C.c_int), ("SL", C.c_int), ("BSL", C.c_int), ("BPC", C.c_int),
        ("offside", C.c_int), ("uniform", C.c_int), ("coeff_kind", C.c_int),
        ("stencil", _dp), ("low", _dp), ("high", _dp), ("coeff", _dp),
    ]


class OrcBC(C.Structure):
    pass


OrcBC._fields_ = [
    ("nl", C.c_int), ("nr", C.c_int), ("a_l", _dp), ("b_l", C.c_double),
    ("a_r", _dp), ("b_r", C.c_double), ("periodic", C.c_int), ("lines", C.POINTER(OrcBC)),
]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        i, d = C.c_int, C.c_double
        ip = C.POINTER(C.c_int)
        llp = C.POINTER(C.c_longlong)
        L.orc_calculate_weights.argtypes = [i, d, _dp, i, _dp]
        L.orc_centered_sizes.argtypes = [i, i, ip, ip, ip]
        L.orc_upwind_sizes.argtypes = [i, i, i, ip, ip, ip]
        L.orc_centered_uniform.argtypes = [i, i, d, i, _dp, _dp, _dp]
        L.orc_centered_nonuniform.argtypes = [i, i, _dp, i, _dp, _dp, _dp]
        L.orc_upwind_uniform.argtypes = [i, i, d, i, i, _dp, _dp, _dp]
        L.orc_upwind_nonuniform.argtypes = [i, i, _dp, i, i, _dp, _dp, _dp]
        L.orc_robin_uniform.argtypes = [d, d, d, d, d, d, d, i, _dp, _dp, _dp, _dp]
        L.orc_robin_nonuniform.argtypes = [d, d, d, d, d, d, _dp, i, i, _dp, _dp, _dp, _dp]
        L.orc_general.argtypes = [_dp, i, _dp, i, _dp, i, i, _dp, _dp, _dp, _dp]
        L.orc_mul_line.argtypes = [C.POINTER(OrcOp), _dp, _dp, i]
        L.orc_mul_line_bpv.argtypes = [C.POINTER(OrcOp), d, d, _dp, _dp, i]
        L.orc_mul_nd.argtypes = [C.POINTER(OrcOp), i, llp, _dp, llp, _dp, i]
        L.orc_ghost_apply.argtypes = [i, C.POINTER(OrcOp), i, llp, C.POINTER(OrcBC), _dp, _dp, i]
        L.orc_composite_mul_padded.argtypes = [i, C.POINTER(OrcOp), i, llp, _dp, _dp]
        L.orc_set_threads.argtypes = [i]
        L.orc_get_threads.restype = i
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with code {rc}")


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def calculate_weights(order: int, x0: float, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x))
    _check(lib().orc_calculate_weights(order, float(x0), _ptr(x), len(x), _ptr(out)), "weights")
    return out


class Op:
    """An oracle DerivativeOperator: FP64 tables built by the C restatement."""

    def __init__(self, wind, d, a, dx, n, axis=0, offside=0, coeff=1.0):
        L = lib()
        SL, BSL, BPC = C.c_int(), C.c_int(), C.c_int()
        self.wind, self.d, self.a, self.len, self.axis = int(wind), d, a, n, axis
        self.offside = offside
        self.uniform = np.ndim(dx) == 0
        if wind:
            _check(L.orc_upwind_sizes(d, a, offside, C.byref(SL), C.byref(BSL), C.byref(BPC)), "sizes")
        else:
            _check(L.orc_centered_sizes(d, a, C.byref(SL), C.byref(BSL), C.byref(BPC)), "sizes")
        self.SL, self.BSL, self.BPC = SL.value, BSL.value, BPC.value
        nin = n - 2 * self.BPC
        if wind:
            ns = self.SL if self.uniform else 2 * nin * self.SL
            nl = self.BPC * self.BSL * (1 if self.uniform else 2)
            nh = (self.BPC + offside) * self.BSL * (1 if self.uniform else 2)
        else:
            ns = self.SL if self.uniform else nin * self.SL
            nl = nh = self.BPC * self.BSL
        self.stencil = np.zeros(max(ns, 1))
        self.low = np.zeros(max(nl, 1))
        self.high = np.zeros(max(nh, 1))
        if self.uniform:
            dx = float(dx)
            f = L.orc_upwind_uniform if wind else L.orc_centered_uniform
            args = (d, a, dx, n, offside) if wind else (d, a, dx, n)
        else:
            self.dxv = np.ascontiguousarray(dx, dtype=np.float64)
            if len(self.dxv) != n + 1:
                raise ValueError("non-uniform dx needs len+1 spacings")
            f = L.orc_upwind_nonuniform if wind else L.orc_centered_nonuniform
            args = (d, a, _ptr(self.dxv), n, offside) if wind else (d, a, _ptr(self.dxv), n)
        _check(f(*args, _ptr(self.stencil), _ptr(self.low), _ptr(self.high)), "constructor")
        if np.ndim(coeff) == 0:
            self.coeff = np.zeros(n) + float(coeff)
            self.coeff_kind = COEFF_NUM_POS if float(coeff) >= 0 else COEFF_NUM_NEG
        else:
            self.coeff = np.zeros(n) + np.asarray(coeff, dtype=np.float64)
            self.coeff_kind = COEFF_VECTOR

    def scaled(self, c):
        """c * A (derivative_operator_functions.jl:165-197): scales the coefficients."""
        import copy
        B = copy.copy(self)
        B.coeff = (np.asarray(c, dtype=np.float64) * self.coeff).astype(np.float64)
        return B

    def struct(self) -> OrcOp:
        return OrcOp(self.axis, self.wind, self.d, self.a, self.len, self.SL, self.BSL,
                     self.BPC, self.offside, int(self.uniform), self.coeff_kind,
                     _ptr(self.stencil), _ptr(self.low), _ptr(self.high), _ptr(self.coeff))

    # 1-D application
    def mul_padded(self, x, out=None, overwrite=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.len) if out is None else out
        s = self.struct()
        _check(lib().orc_mul_line(C.byref(s), _ptr(x), _ptr(out), int(overwrite)), "mul_line")
        return out

    def mul_bpv(self, l, r, u, out=None, overwrite=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(self.len) if out is None else out
        s = self.struct()
        _check(lib().orc_mul_line_bpv(C.byref(s), float(l), float(r), _ptr(u), _ptr(out),
                                      int(overwrite)), "mul_bpv")
        return out

    def matrix(self, bpv=False):
        """len x (len+2) matrix by multiplying unit vectors (the reference tests'
        convert_by_multiplication, derivative_operators_interface.jl:102-117)."""
        n = self.len
        A = np.zeros((n, n + 2))
        for j in range(n + 2):
            v = np.zeros(n + 2)
            v[j] = 1.0
            if bpv:
                A[:, j] = self.mul_bpv(v[0], v[-1], v[1:-1])
            else:
                A[:, j] = self.mul_padded(v)
        return A


class BC:
    """An affine (Robin-family) BC: ghost = a . u[1:k] + b on each side."""

    def __init__(self, a_l, b_l, a_r, b_r, periodic=0, lines=None):
        self.a_l = np.ascontiguousarray(a_l, dtype=np.float64)
        self.a_r = np.ascontiguousarray(a_r, dtype=np.float64)
        self.b_l, self.b_r = float(b_l), float(b_r)
        self.periodic = int(periodic)   # 1: line wrap, 2: N-D periodic padding (first/last)
        self.lines = lines              # N-D: one BC per line (column-major over the other axes)

    def struct(self) -> OrcBC:
        s = OrcBC(len(self.a_l), len(self.a_r), _ptr(self.a_l), self.b_l, _ptr(self.a_r), self.b_r,
                  int(self.periodic), None)
        if self.lines is not None:
            self._line_structs = (OrcBC * len(self.lines))(*[b.struct() for b in self.lines])
            s.lines = C.cast(self._line_structs, C.POINTER(OrcBC))
        return s

    @staticmethod
    def per_line(bcs):
        """One BC per line (a list in column-major order over the other axes)."""
        return BC([], 0.0, [], 0.0, lines=list(bcs))

    def pad(self, u):
        u = np.asarray(u, dtype=np.float64)
        if self.periodic == 1:
            return float(u[-1]), float(u[0])
        if self.periodic == 2:
            return float(u[0]), float(u[-1])
        gl = float(np.dot(self.a_l, u[: len(self.a_l)]) + self.b_l)
        gr = float(np.dot(self.a_r, u[len(u) - len(self.a_r):]) + self.b_r)
        return gl, gr


def robin(l, r, dx, order=1) -> BC:
    """RobinBC(l, r, dx, order): bc_operators.jl:26-60."""
    al, bl, cl = map(float, l)
    ar, br, cr = map(float, r)
    a_l = np.zeros(order)
    a_r = np.zeros(order)
    b_l = C.c_double()
    b_r = C.c_double()
    if np.ndim(dx) == 0:
        _check(lib().orc_robin_uniform(al, bl, cl, ar, br, cr, float(dx), order, _ptr(a_l),
                                       C.byref(b_l), _ptr(a_r), C.byref(b_r)), "robin")
    else:
        dxv = np.ascontiguousarray(dx, dtype=np.float64)
        _check(lib().orc_robin_nonuniform(al, bl, cl, ar, br, cr, _ptr(dxv), len(dxv), order,
                                          _ptr(a_l), C.byref(b_l), _ptr(a_r), C.byref(b_r)), "robin")
    return BC(a_l, b_l.value, a_r, b_r.value)


def dirichlet(vl, vr) -> BC:
    """DirichletBC(αl, αr) = RobinBC((1,0,αl),(1,0,αr), 1.0, 2): bc_operators.jl:176-178."""
    return robin((1.0, 0.0, vl), (1.0, 0.0, vr), 1.0, 2)


def dirichlet0() -> BC:
    return dirichlet(0.0, 0.0)


def general(alpha_l, alpha_r, dx, order=1) -> BC:
    """GeneralBC(αl, αr, dx, order): bc_operators.jl:85-170."""
    al = np.ascontiguousarray(alpha_l, dtype=np.float64)
    ar = np.ascontiguousarray(alpha_r, dtype=np.float64)
    a_l = np.zeros(max(order + len(al) - 3, 0))
    a_r = np.zeros(max(order + len(ar) - 3, 0))
    b_l = C.c_double()
    b_r = C.c_double()
    if np.ndim(dx) == 0:
        dxv, ndx = np.array([float(dx)]), 0
    else:
        dxv = np.ascontiguousarray(dx, dtype=np.float64)
        ndx = len(dxv)
    _check(lib().orc_general(_ptr(al), len(al), _ptr(ar), len(ar), _ptr(dxv), ndx, order,
                             _ptr(a_l), C.byref(b_l), _ptr(a_r), C.byref(b_r)), "general")
    return BC(a_l, b_l.value, a_r, b_r.value)


def periodic() -> BC:
    """PeriodicBC(T) applied to a line: ghosts u[end], u[1] (bc_operators.jl:192)."""
    return BC([], 0.0, [], 0.0, periodic=1)


def periodic_nd() -> BC:
    """PeriodicBC padding of an N-D array through MultiDimDirectionalBC{T,PeriodicBC{T}} /
    ComposedMultiDimBC{T,PeriodicBC{T}}: ghosts u[1], u[end] (multi_dim_bc_operators.jl:
    221-252)."""
    return BC([], 0.0, [], 0.0, periodic=2)


def neumann(g, dx, order=1) -> BC:
    """NeumannBC((γl, γr), dx, order) = RobinBC((0,1,γl),(0,1,γr), dx, order): :173-175."""
    return robin((0.0, 1.0, g[0]), (0.0, 1.0, g[1]), dx, order)


def _ll(a):
    arr = (C.c_longlong * len(a))(*[int(v) for v in a])
    return arr


def ghost_apply(ops, bcs, u, bpv_1d=True) -> np.ndarray:
    """sum_t (L_t * Q) * u with Q the composed per-axis BCs; u column-major (numpy shape
    reversed from the Julia size)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    size = list(u.shape[::-1])
    ndim = len(size)
    out = np.zeros_like(u)
    structs = (OrcOp * len(ops))(*[o.struct() for o in ops])
    bstructs = (OrcBC * ndim)(*[b.struct() for b in bcs])
    _check(lib().orc_ghost_apply(len(ops), structs, ndim, _ll(size), bstructs, _ptr(u),
                                 _ptr(out), int(bpv_1d)), "ghost_apply")
    return out


def mul_nd(op, M, out_shape, out=None, overwrite=True) -> np.ndarray:
    """_mul!(out, L, M): out_shape / M in numpy (reversed) order."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    out = np.zeros(out_shape) if out is None else out
    s = op.struct()
    osz = list(out.shape[::-1])
    msz = list(M.shape[::-1])
    _check(lib().orc_mul_nd(C.byref(s), len(osz), _ll(osz), _ptr(out), _ll(msz), _ptr(M),
                            int(overwrite)), "mul_nd")
    return out


def composite_mul_padded(ops, M) -> np.ndarray:
    M = np.ascontiguousarray(M, dtype=np.float64)
    out_shape = tuple(s - 2 for s in M.shape)
    out = np.zeros(out_shape)
    structs = (OrcOp * len(ops))(*[o.struct() for o in ops])
    osz = list(out_shape[::-1])
    _check(lib().orc_composite_mul_padded(len(ops), structs, len(osz), _ll(osz), _ptr(M),
                                          _ptr(out)), "composite")
    return out


# ------------------------------------------------------------------------------------------
# Adaptive Runge-Kutta time stepping (the reference's method-of-lines tests integrate
# `(u, p, t) -> A*bc*u` with OrdinaryDiffEq's Tsit5, heat_eqn.jl:4-60; OrdinaryDiffEq is an
# external dependency, absent here and unpinned: no Manifest.toml).  Restated from the
# published method: the 5(4) pair of Tsitouras (2011, Comput. Math. Appl. 62: 770-775) and the
# DOPRI5 step-size controller of Hairer, Norsett & Wanner (Solving ODEs I, 2nd ed., II.4).

TSIT5_C = np.array([0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0])
TSIT5_A = np.array([
    [0, 0, 0, 0, 0, 0],
    [0.161, 0, 0, 0, 0, 0],
    [-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0],
    [2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0],
    [5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0],
    [5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383, 0],
    [0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774],
], dtype=np.float64)
TSIT5_B = np.append(TSIT5_A[6], 0.0)          # first same as last: b = the last row of A
TSIT5_BTILDE = np.array([-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                         -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                         0.015151515151515152])


def _rms(x):
    return float(np.sqrt(np.mean(np.square(x))))


def tsit5_initial_dt(f, u0, f0, t0, t1, atol, rtol):
    """Starting step size (Hairer, Norsett & Wanner, II.4) with the scaled RMS norm."""
    sc = atol + rtol * np.abs(u0)
    d0, d1 = _rms(u0 / sc), _rms(f0 / sc)
    h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    h0 = min(h0, t1 - t0)
    f1 = f(u0 + h0 * f0)
    d2 = _rms((f1 - f0) / sc) / h0
    m = max(d1, d2)
    h1 = max(1e-6, h0 * 1e-3) if m <= 1e-15 else (0.01 / m) ** (1.0 / 5.0)
    return min(100 * h0, h1)


def tsit5(f, u0, t0, t1, dt0=None, atol=1e-6, rtol=1e-3, max_attempts=10 ** 6):
    """Integrate du/dt = f(u) from t0 to t1 (steps clamped to end at t1).  Returns
    (u, t, naccept, nreject)."""
    u = np.array(u0, dtype=np.float64)
    k = [None] * 7
    k[0] = f(u)
    dt = dt0 if (dt0 is not None and dt0 > 0) else tsit5_initial_dt(f, u, k[0], t0, t1, atol, rtol)
    dt = min(dt, t1 - t0)
    t, facold, rejected = t0, 1e-4, False
    na = nr = 0
    beta, safe = 0.04, 0.9
    for _ in range(max_attempts):
        if t1 - t <= 1e-14 * max(abs(t1), 1.0):
            break
        for s in range(1, 7):
            acc = TSIT5_A[s, 0] * k[0]
            for j in range(1, s):
                acc = acc + TSIT5_A[s, j] * k[j]
            k[s] = f(u + dt * acc)
        unew = u + dt * (TSIT5_A[6, 0] * k[0] + sum(TSIT5_A[6, j] * k[j] for j in range(1, 6)))
        k[6] = f(unew)
        e = dt * sum(TSIT5_BTILDE[j] * k[j] for j in range(7))
        sc = atol + rtol * np.maximum(np.abs(u), np.abs(unew))
        err = _rms(e / sc)
        fac11 = err ** (0.2 - beta * 0.75)
        if err <= 1.0:
            t += dt
            na += 1
            fac = max(0.1, min(5.0, fac11 / facold ** beta / safe))
            dtn = dt / fac
            if rejected:
                dtn = min(dtn, dt)
            facold = max(err, 1e-4)
            rejected = False
            u = unew
            k[0] = k[6]
        else:
            nr += 1
            dtn = dt / min(5.0, fac11 / safe)
            rejected = True
        dt = min(dtn, t1 - t)
    return u, t, na, nr


# ------------------------------------------------------------------------------------------
# Vector calculus (vector_calculus_operators.jl:5-88, vector_calculus_convolutions.jl).  Every
# derivative is the centered line operator along one axis of a padded array; for 2-D and 3-D
# arrays the reference's kernels (vector_calculus_convolutions.jl:173-1067) index the
# transverse axes of the padded input at 1:n, not at the interior 2:n+1 (restated as read);
# 1-D has no transverse axis.

def vc_line_op(op, M, k, transverse_view):
    """op (a line operator along Julia axis k, 0-based) applied to every line of the padded
    array M (numpy order), transverse axes at 0..n-1 (transverse_view) or 1..n."""
    nd = M.ndim
    ax = nd - 1 - k
    Mt = np.moveaxis(np.asarray(M, dtype=np.float64), ax, -1)
    tshape = tuple(s - 2 for s in Mt.shape[:-1])
    off = 0 if transverse_view else 1
    out = np.zeros(tshape + (Mt.shape[-1] - 2,))
    for idx in np.ndindex(*tshape):
        out[idx] = op.mul_padded(np.ascontiguousarray(Mt[tuple(i + off for i in idx)]))
    return np.moveaxis(out, -1, ax)


def gradient(ops, M):
    """Gradient * M: shape (N,) + interior shape (component first in numpy order)."""
    nd = M.ndim
    return np.stack([vc_line_op(ops[k], M, k, nd in (2, 3)) for k in range(nd)])


def divergence(ops, M):
    """Divergence * M, M of shape (N,) + padded shape."""
    nd = M.ndim - 1
    out = None
    for k in range(nd):
        v = vc_line_op(ops[k], M[k], k, nd in (2, 3))
        out = v if out is None else out + v
    return out


def curl(ops, M):
    """Curl * M, M of shape (3,) + padded 3-D shape: (d2 u3 - d3 u2, d3 u1 - d1 u3,
    d1 u2 - d2 u1) with d_k along Julia axis k."""
    d = lambda k, c: vc_line_op(ops[k], M[c], k, True)   # noqa: E731
    return np.stack([d(1, 2) - d(2, 1), d(2, 0) - d(0, 2), d(0, 1) - d(1, 0)])


def nonlinear_diffusion(s, f, approx, p, q, dx, n):
    """nonlinear_diffusion!{N}(du, s, f, approx, p, q, dx, n) (derivative_operator.jl:34-66),
    restated term by term with the oracle's line kernel: D^k = CenteredDifference(k, approx,
    dx, n) on the padded vectors, D^0 = the interior."""
    def D(k, v):
        if k == 0:
            return np.asarray(v, dtype=np.float64)[1:n + 1]
        return Op(0, k, approx, dx, n).mul_padded(np.ascontiguousarray(v, dtype=np.float64))
    du = D(f, q) * D(s, p)
    for l in range(1, s):
        du = du + (math.comb(s, l) * D(l + f, q)) * D(s - l, p)
    return du + D(f + s, q) * D(0, p)
