"""GPU test of dop_plan_csr_rows_noalloc, the row-pointer export that allocates nothing (its
scan partial sums and bandwidth counters live in the library's static device scratch).

Through the C ABI: on two calls in a row, the second on a smaller plan (so the scratch holds
the first call's contents), row_ptr, nnz and the bandwidths equal dop_plan_csr_rows' bit for
bit, and the row pointers it writes drive dop_plan_csr_fill to the same columns, values and
affine part.  A plan beyond the scratch's maximum (2^28 rows) is rejected with DOP_EBUDGET
before anything is launched."""
import ctypes as C

import numpy as np
import pytest
import torch

import diffeqoperators_amd as D
from diffeqoperators_amd import _abi

pytestmark = pytest.mark.gpu


def _lib():
    L = _abi.lib()
    i64p = C.POINTER(C.c_int64)
    L.dop_plan_csr_rows_noalloc.argtypes = [C.c_void_p, C.c_void_p, i64p, i64p, i64p, C.c_void_p]
    return L


def _laplacian2d(n1, n2, rng):
    h = 1.0 / (n1 + 1)
    dx = h * (1 + 0.5 * rng.random(n2 + 1))
    A = (D.CenteredDifference[1](2, 4, h, n1) + D.CenteredDifference[2](2, 6, dx, n2) +
         D.UpwindDifference[1](1, 3, h, n1, -rng.standard_normal(n1))) * \
        D.compose(D.MultiDimBC[1](D.RobinBC((1.0, 0.5, 0.2), (1.0, 0.4, -0.3), h, 2), (n1, n2)),
                  D.MultiDimBC[2](D.DirichletBC(0.5, -1.0), (n1, n2)))
    return D.plan_for_shape(A, (n2, n1), torch.float64)


def _export(L, plan, noalloc):
    dev = torch.device("cuda")
    n = plan.out_elems
    rp = torch.full((n + 1,), -7, dtype=torch.int64, device=dev)   # row_ptr[0] must be written
    nnz, kl, ku = C.c_int64(), C.c_int64(), C.c_int64()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    f = L.dop_plan_csr_rows_noalloc if noalloc else L.dop_plan_csr_rows
    _abi.check(f(plan.handle, C.c_void_p(rp.data_ptr()), C.byref(nnz), C.byref(kl), C.byref(ku), s),
               "csr rows")
    col = torch.empty(max(nnz.value, 1), dtype=torch.int64, device=dev)
    val = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=dev)
    b = torch.empty(n, dtype=torch.float64, device=dev)
    _abi.check(L.dop_plan_csr_fill(plan.handle, C.c_void_p(rp.data_ptr()), C.c_void_p(col.data_ptr()),
                                   C.c_void_p(val.data_ptr()), C.c_void_p(b.data_ptr()), s), "csr fill")
    torch.cuda.synchronize()
    return (rp.cpu().numpy(), nnz.value, kl.value, ku.value, col[: nnz.value].cpu().numpy(),
            val[: nnz.value].cpu().numpy(), b.cpu().numpy())


def test_noalloc_rows_equal_allocating_export_twice_shrinking():
    L = _lib()
    rng = np.random.default_rng(77)
    # 600 x 700 = 420000 rows (103 scan tiles), then 90 x 50 (2 tiles): the second call runs on
    # a scratch that still holds the first call's partial sums and bandwidths
    for n1, n2 in ((600, 700), (90, 50)):
        print(f"[no_alloc] plan {n1}x{n2}", flush=True)
        plan = _laplacian2d(n1, n2, rng)
        ref = _export(L, plan, noalloc=False)
        got = _export(L, plan, noalloc=True)
        assert got[0][0] == 0
        np.testing.assert_array_equal(got[0], ref[0])
        assert got[1:4] == ref[1:4]
        np.testing.assert_array_equal(got[4], ref[4])
        np.testing.assert_array_equal(got[5].view(np.int64), ref[5].view(np.int64))
        np.testing.assert_array_equal(got[6].view(np.int64), ref[6].view(np.int64))
        assert ref[1] > 0 and ref[2] > 0 and ref[3] > 0


def test_noalloc_rejects_plans_beyond_the_scratch():
    L = _lib()
    n = 1024
    A = (D.CenteredDifference[1](2, 2, 1.0, n) + D.CenteredDifference[2](2, 2, 1.0, n) +
         D.CenteredDifference[3](2, 2, 1.0, 2 * n)) * D.compose(*D.Dirichlet0BC(float, (n, n, 2 * n)))
    plan = D.plan_for_shape(A, (2 * n, n, n), torch.float64)    # 2^31 rows > 2^28
    rp = torch.empty(8, dtype=torch.int64, device="cuda")   # never written: rejected first
    nnz, kl, ku = C.c_int64(-1), C.c_int64(), C.c_int64()
    rc = L.dop_plan_csr_rows_noalloc(plan.handle, C.c_void_p(rp.data_ptr()), C.byref(nnz),
                                     C.byref(kl), C.byref(ku), None)
    assert rc == _abi.EBUDGET
    assert b"static scratch" in L.dop_last_error()
    assert nnz.value == -1
