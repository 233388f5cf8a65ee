"""The device canonical-CSR check (rvsp_csr_check_device) against the
oracle's restatement of the reference's rvsp_csr_check and, where it was
built, the reference library itself; and the one-shot path, which runs the
device check on its staged operands."""
import ctypes as C
import time

import numpy as np
import pytest

import cases
import rvsp
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import Csr, csr_check_device, device_alloc, device_free, to_device

pytestmark = pytest.mark.gpu
CASES = cases.csr_check_cases()
L = F.lib()


def on_device(rows, cols, rp, ci, nnz=None):
    """Uploads the arrays, runs the device check, frees them."""
    nnz = (0 if ci is None else len(ci)) if nnz is None else nnz
    d_rp = device_alloc(4 * len(rp))
    to_device(d_rp, np.asarray(rp, np.int32))
    d_ci = None
    if ci is not None:
        d_ci = device_alloc(4 * max(len(ci), 1))
        to_device(d_ci, np.asarray(ci, np.int32))
    try:
        return csr_check_device(rows, cols, nnz, d_rp, d_ci)
    finally:
        device_free(d_rp)
        device_free(d_ci)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_check_matches_oracle(gpu, case):
    _, rows, cols, rp, ci = case
    want = oracle.csr_check(rows, cols, rp, ci)
    assert on_device(rows, cols, rp, ci) == want
    if oracle.ref_available():
        assert oracle.ref_csr_check(rows, cols, rp, ci) == want


def test_device_check_argument_cases(gpu):
    st, bad = C.c_int32(), C.c_int32()
    assert L.rvsp_csr_check_device(3, 4, 0, None, None, C.byref(st), C.byref(bad)) == F.SUCCESS
    assert (st.value, bad.value) == (oracle.CSR_BAD_ROW_PTR, -1)
    assert L.rvsp_csr_check_device(-1, 4, 0, None, None, C.byref(st), None) == F.SUCCESS
    assert st.value == oracle.CSR_BAD_ROW_PTR
    assert L.rvsp_csr_check_device(1, 4, 0, None, None, None, None) == F.ERROR_NULL_POINTER


def test_row_pointer_past_nnz_is_malformed(gpu):
    # The device check knows how many column entries exist: a row ending
    # past them is malformed at that row (a host check would read past the
    # array instead). Earlier failures still win.
    rp = np.array([0, 2, 3, 9, 9], np.int32)
    ci = np.array([0, 2, 1, 0, 3], np.int32)
    assert on_device(4, 4, rp, ci) == (oracle.CSR_BAD_ROW_PTR, 2)
    ci_dup = np.array([1, 1, 1, 0, 3], np.int32)
    assert on_device(4, 4, rp, ci_dup) == (oracle.CSR_DUPLICATE, 0)


def test_full_size_laplacian(gpu):
    # 1,048,576 rows, 5,238,784 entries: valid, then a fault near the end and
    # another near the start (the earlier one wins); times both checks.
    A = gen.laplacian_2d(1024)
    rp = np.asarray(A.row_ptr, np.int32)
    ci = np.asarray(A.col_idx, np.int32)
    n, rows = len(ci), A.rows
    d_rp, d_ci = device_alloc(4 * (rows + 1)), device_alloc(4 * n)
    try:
        to_device(d_rp, rp)
        to_device(d_ci, ci)
        assert csr_check_device(rows, A.cols, n, d_rp, d_ci) == (oracle.CSR_OK, -1)
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            csr_check_device(rows, A.cols, n, d_rp, d_ci)
        t_dev = (time.perf_counter() - t0) / reps
        late = rows - 3
        ci2 = ci.copy()
        p = int(rp[late])
        ci2[p], ci2[p + 1] = ci2[p + 1], ci2[p]
        want = oracle.csr_check(rows, A.cols, rp, ci2)
        assert want == (oracle.CSR_UNSORTED, late)
        to_device(d_ci, ci2)
        assert csr_check_device(rows, A.cols, n, d_rp, d_ci) == want
        early = 5
        ci2[int(rp[early]) + 1] = ci2[int(rp[early])]
        want = oracle.csr_check(rows, A.cols, rp, ci2)
        assert want == (oracle.CSR_DUPLICATE, early)
        to_device(d_ci, ci2)
        assert csr_check_device(rows, A.cols, n, d_rp, d_ci) == want
    finally:
        device_free(d_rp)
        device_free(d_ci)
    host = np.ascontiguousarray(ci)
    t0 = time.perf_counter()
    L.rvsp_csr_check(rows, A.cols, rp.ctypes.data, host.ctypes.data, None)
    t_host = time.perf_counter() - t0
    print(f"\ncanonical check, 5-point Laplacian 1024^2 ({n} entries): device {t_dev * 1e3:.3f} ms "
          f"per call (synchronous), host {t_host * 1e3:.3f} ms")


def test_oneshot_rejects_invalid_operands_on_the_device(gpu):
    good = Csr(3, 3, [0, 2, 3, 3], [0, 2, 1], [1, 2, 3])
    unsorted = Csr(3, 3, [0, 2, 3, 3], [2, 0, 1], [1, 2, 3])
    out_of_range = Csr(3, 3, [0, 2, 3, 3], [0, 3, 1], [1, 2, 3])
    short = Csr(3, 3, [0, 2, 3, 3], [0, 2, 1], [1, 2, 3])
    past_nnz = short.struct()
    for A, B in ((unsorted, good), (good, unsorted), (out_of_range, good)):
        Cm = F.CsrMatrix()
        for backend in (F.BACKEND_SCALAR, F.BACKEND_HIP):
            opts = F.SpgemmOptions(backend, F.DTYPE_FP32, F.DTYPE_FP32)
            assert L.rvsp_spgemm_csr(C.byref(A.struct()), C.byref(B.struct()), C.byref(Cm),
                                     C.byref(opts)) == F.ERROR_INVALID_CSR
    # The struct claims fewer entries than its row pointers cover.
    past_nnz.nnz = 2
    Cm = F.CsrMatrix()
    assert L.rvsp_spgemm_csr(C.byref(past_nnz), C.byref(good.struct()), C.byref(Cm),
                             None) == F.ERROR_INVALID_CSR
    # Valid operands still multiply, bit-exact against the oracle.
    got = rvsp.spgemm_oneshot(good, good)
    rp, ci, val = oracle.spgemm(good, good)
    assert got.row_ptr.tobytes() == rp.tobytes() and got.col_idx.tobytes() == ci.tobytes()
    assert got.values.tobytes() == val.tobytes()
