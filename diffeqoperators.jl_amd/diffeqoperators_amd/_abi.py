"""ctypes view of include/dop.h and the loader of libdop.so.

The library is required: there is no CPU fallback.  Loading fails loudly when the built
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("DOP_LIBRARY", os.path.join(_PKG_ROOT, "lib", "libdop.so"))

ABI_VERSION = 2

OK, EINVAL, EDIM, EHIP, ENOMEM, EUNSUPPORTED, EBUDGET, ESINGULAR = 0, -1, -2, -3, -4, -5, -6, -7
F64, F32, BF16 = 0, 1, 2
ACC_AUTO, ACC_F64, ACC_F32, ACC_F32_KAHAN, ACC_F32_WIDE = 0, 1, 2, 3, 4
INPUT_GHOST, INPUT_PADDED = 0, 1
SEM_LINES, SEM_BPV = 0, 1
COEFF_VECTOR, COEFF_NUM_POS, COEFF_NUM_NEG = 0, 1, 2
MAX_BC_TAPS = 8
BC_AFFINE, BC_PERIODIC, BC_PERIODIC_ND = 0, 1, 2

_dp = C.POINTER(C.c_double)


class DopOperator(C.Structure):
    _fields_ = [
        ("axis", C.c_int32), ("wind", C.c_int32), ("derivative_order", C.c_int32),
        ("approximation_order", C.c_int32), ("len", C.c_int32), ("stencil_length", C.c_int32),
        ("boundary_stencil_length", C.c_int32), ("boundary_point_count", C.c_int32),
        ("offside", C.c_int32), ("uniform_grid", C.c_int32), ("coeff_kind", C.c_int32),
        ("reserved", C.c_int32),
        ("stencil_coefs", _dp), ("low_boundary_coefs", _dp), ("high_boundary_coefs", _dp),
        ("coefficients", _dp),
    ]


class DopAffineBC(C.Structure):
    _fields_ = [
        ("present", C.c_int32), ("n_l", C.c_int32), ("n_r", C.c_int32), ("kind", C.c_int32),
        ("a_l", C.c_double * MAX_BC_TAPS), ("b_l", C.c_double),
        ("a_r", C.c_double * MAX_BC_TAPS), ("b_r", C.c_double),
    ]


class DopDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("storage", C.c_int32), ("accum", C.c_int32), ("input", C.c_int32),
        ("semantics", C.c_int32), ("n_ops", C.c_int32), ("size", C.c_int64 * 3),
        ("padded", C.c_int32 * 3), ("transverse_view", C.c_int32),
        ("ops", C.POINTER(DopOperator)), ("bc", DopAffineBC * 3),
        ("line_bc", C.POINTER(DopAffineBC) * 3),
    ]


class DopError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libdop error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded libdop.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libdop.so not found at {LIB_PATH}: build it with "
            "`make -C diffeqoperators.jl_amd/csrc` (the operators have no CPU path)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i = C.c_int
    L.dop_version.restype = i
    L.dop_last_error.restype = C.c_char_p
    L.dop_plan_create.argtypes = [C.POINTER(DopDesc), C.POINTER(vp)]
    L.dop_plan_destroy.argtypes = [vp]
    L.dop_plan_destroy.restype = None
    L.dop_plan_input_elems.argtypes = [vp]
    L.dop_plan_input_elems.restype = C.c_int64
    L.dop_plan_output_elems.argtypes = [vp]
    L.dop_plan_output_elems.restype = C.c_int64
    L.dop_apply.argtypes = [vp, vp, vp, i, vp]
    L.dop_plan_autotune.argtypes = [vp, vp, C.POINTER(i)]
    L.dop_plan_update.argtypes = [vp, C.POINTER(DopDesc), vp]
    L.dop_plan_variant.argtypes = [vp, C.POINTER(i)]
    L.dop_batch_create.argtypes = [i, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.dop_batch_apply.argtypes = [vp, i, vp]
    L.dop_batch_autotune.argtypes = [vp, vp, C.POINTER(i)]
    L.dop_rk4_chunk.argtypes = [vp]
    L.dop_batch_destroy.argtypes = [vp]
    L.dop_batch_destroy.restype = None
    L.dop_rk4_create.argtypes = [i, C.POINTER(vp), _dp, _dp, C.POINTER(vp), vp, C.POINTER(vp)]
    L.dop_rk4_step.argtypes = [vp, i, vp]
    L.dop_rk4_destroy.argtypes = [vp]
    L.dop_rk4_destroy.restype = None
    L.dop_server_create.argtypes = [C.c_int64, i, i, vp, C.POINTER(vp)]
    L.dop_server_submit.argtypes = [vp, vp, C.c_double, C.c_double, i, vp, vp, C.c_double,
                                    C.POINTER(C.c_int64)]
    L.dop_server_run.argtypes = [vp]
    L.dop_server_set_open.argtypes = [vp, i]
    L.dop_server_now.argtypes = [vp]
    L.dop_server_now.restype = C.c_double
    L.dop_server_times.argtypes = [vp, C.c_int64, _dp, _dp, _dp]
    L.dop_server_stats.argtypes = [vp, C.POINTER(i), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.dop_server_destroy.argtypes = [vp]
    L.dop_server_destroy.restype = None
    i64p = C.POINTER(C.c_int64)
    L.dop_plan_csr_rows.argtypes = [vp, vp, i64p, i64p, i64p, vp]
    L.dop_plan_csr_fill.argtypes = [vp, vp, vp, vp, vp, vp]
    L.dop_solver_create.argtypes = [vp, vp, C.POINTER(vp)]
    L.dop_solver_solve.argtypes = [vp, vp, vp, i, vp]
    L.dop_solver_band.argtypes = [vp, i64p, i64p, i64p]
    L.dop_solver_destroy.argtypes = [vp]
    L.dop_solver_destroy.restype = None
    dbl = C.c_double
    L.dop_tsit5_create.argtypes = [vp, dbl, vp, dbl, dbl, dbl, dbl, dbl, vp, C.POINTER(vp)]
    L.dop_tsit5_run.argtypes = [vp, C.c_int64, i, vp]
    L.dop_tsit5_set_end.argtypes = [vp, dbl, vp]
    L.dop_tsit5_status.argtypes = [vp, _dp, _dp, i64p, i64p, C.POINTER(i), vp]
    L.dop_tsit5_destroy.argtypes = [vp]
    L.dop_tsit5_destroy.restype = None
    L.dop_nonlinear_diffusion.argtypes = [C.POINTER(vp), C.POINTER(vp), i, vp, vp, vp, vp]
    if L.dop_version() != ABI_VERSION:
        raise RuntimeError(f"libdop ABI {L.dop_version()} != {ABI_VERSION}")
    _lib = L
    return L


def check(rc, what=""):
    if rc != OK:
        msg = lib().dop_last_error().decode(errors="replace")
        raise DopError(rc, f"{what}: {msg}" if what else msg)
