"""rvsp_csr_check_device_noalloc, the device canonical-CSR check whose
scratch is a static device array, against rvsp_csr_check_device (which
takes its scratch from the library's buffer cache): same verdicts and rows
on every planted-fault structure, on two calls in a row, the second with a
smaller input; and no growth of the library's device buffer cache."""
import ctypes as C

import numpy as np
import pytest

import cases
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import device_alloc, device_free, to_device

pytestmark = pytest.mark.gpu
L = F.lib()


def upload(rp, ci):
    d_rp = device_alloc(4 * max(len(rp), 1))
    to_device(d_rp, np.asarray(rp, np.int32))
    d_ci = None
    if ci is not None:
        d_ci = device_alloc(4 * max(len(ci), 1))
        to_device(d_ci, np.asarray(ci, np.int32))
    return d_rp, d_ci


def check(fn, rows, cols, nnz, d_rp, d_ci):
    st, bad = C.c_int32(-99), C.c_int32(-99)
    rc = fn(rows, cols, nnz, C.c_void_p(d_rp), C.c_void_p(d_ci) if d_ci else None,
            C.byref(st), C.byref(bad))
    assert rc == F.SUCCESS
    return st.value, bad.value


def both(rows, cols, rp, ci):
    nnz = 0 if ci is None else len(ci)
    d_rp, d_ci = upload(rp, ci)
    try:
        want = check(L.rvsp_csr_check_device, rows, cols, nnz, d_rp, d_ci)
        got = check(L.rvsp_csr_check_device_noalloc, rows, cols, nnz, d_rp, d_ci)
    finally:
        device_free(d_rp)
        device_free(d_ci)
    return got, want


def test_noalloc_matches_check_device_on_every_case(gpu):
    for name, rows, cols, rp, ci in cases.csr_check_cases():
        got, want = both(rows, cols, rp, ci)
        assert got == want, name
        assert got == oracle.csr_check(rows, cols, rp, ci), name


def test_two_calls_second_smaller(gpu):
    # A large faulty structure first (its failure leaves a small key in the
    # scratch), then a smaller valid one and a smaller faulty one: each
    # call's verdict must be its own.
    A = gen.laplacian_2d(512)
    rp = np.asarray(A.row_ptr, np.int32)
    ci = np.asarray(A.col_idx, np.int32).copy()
    ci[1000] = ci[999]  # duplicate early in the matrix
    got, want = both(A.rows, A.cols, rp, ci)
    assert got == want and got[0] == oracle.CSR_DUPLICATE
    B = gen.uniform_k(300, 7, seed=4)
    got, want = both(B.rows, B.cols, B.row_ptr, B.col_idx)
    assert got == want == (oracle.CSR_OK, -1)
    ci2 = np.asarray(B.col_idx, np.int32).copy()
    ci2[-1] = B.cols  # out of range in the last row
    got, want = both(B.rows, B.cols, B.row_ptr, ci2)
    assert got == want == (oracle.CSR_COL_OUT_OF_RANGE, B.rows - 1)


def test_argument_cases_and_no_cache_growth(gpu):
    st, bad = C.c_int32(), C.c_int32()
    assert L.rvsp_csr_check_device_noalloc(1, 4, 0, None, None, None, None) == \
        F.ERROR_NULL_POINTER
    assert L.rvsp_csr_check_device_noalloc(3, 4, 0, None, None, C.byref(st),
                                           C.byref(bad)) == F.SUCCESS
    assert (st.value, bad.value) == (oracle.CSR_BAD_ROW_PTR, -1)
    B = gen.uniform_k(2000, 9, seed=8)
    d_rp, d_ci = upload(B.row_ptr, B.col_idx)
    try:
        assert L.rvsp_hip_pool_release() == F.SUCCESS
        for _ in range(3):
            assert check(L.rvsp_csr_check_device_noalloc, B.rows, B.cols, B.nnz, d_rp,
                         d_ci) == (oracle.CSR_OK, -1)
        assert L.rvsp_hip_pool_cached_bytes() == 0
    finally:
        device_free(d_rp)
        device_free(d_ci)
