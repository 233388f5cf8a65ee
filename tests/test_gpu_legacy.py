"""FP64 -> FP64 and INT8 -> INT32 products through rvsp_spgemm_csr on the
GPU, bit for bit against the oracle's restatement of the reference's legacy
kernels (and the reference library itself where it was built); the status
codes of each legacy wrapper."""
import ctypes as C

import numpy as np
import pytest

import cases
import rvsp
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import Csr

pytestmark = pytest.mark.gpu
L = F.lib()


def same(got, want):
    rp, ci, val = want
    assert got.row_ptr.tobytes() == rp.tobytes(), "row_ptr differs"
    assert got.col_idx.tobytes() == ci.tobytes(), "col_idx differs"
    assert got.values.dtype == val.dtype
    if got.values.tobytes() != val.tobytes():
        bad = np.flatnonzero(got.values != val)
        raise AssertionError(f"{bad.size} values differ, first at {bad[:5]}")


def status(A, B, dt_in, dt_out, backend=F.BACKEND_SCALAR):
    Cm = F.CsrMatrix()
    opts = F.SpgemmOptions(backend, dt_in, dt_out)
    st = L.rvsp_spgemm_csr(C.byref(A.struct()), C.byref(B.struct()), C.byref(Cm), C.byref(opts))
    if st == F.SUCCESS:
        L.rvsp_csr_destroy(C.byref(Cm))
    return st


@pytest.mark.parametrize("i8", [False, True], ids=["f64", "i8"])
@pytest.mark.parametrize("seed", range(6))
def test_legacy_products_match_oracle(gpu, i8, seed):
    A, B = cases.legacy_pair(seed, i8)
    want = oracle.spgemm_legacy(A, B)
    for backend in (F.BACKEND_SCALAR, F.BACKEND_HIP):
        same(rvsp.spgemm_legacy(A, B, backend=backend), want)
    if oracle.ref_available():
        st, ref = oracle.ref_spgemm_legacy(A, B)
        assert st == 0
        same(rvsp.spgemm_legacy(A, B), ref)


@pytest.mark.parametrize("i8", [False, True], ids=["f64", "i8"])
def test_legacy_wide_rows_and_many_chunks(gpu, i8):
    # C rows of several thousand columns: the numeric kernel takes them in
    # chunks of positions and narrows every B row to each chunk.
    A, B = cases.legacy_pair(42, i8, rows=64, inner=400, cols=60000, mean=6.0, long_rows=8)
    want = oracle.spgemm_legacy(A, B)
    assert int(np.diff(want[0]).max()) > 3000
    same(rvsp.spgemm_legacy(A, B), want)


def test_legacy_fp64_full_size_laplacian(gpu):
    # 1,048,576 rows, values drawn from {-1, 1} and fractions: cancellations
    # and order-sensitive sums in one product.
    A = gen.laplacian_2d(1024)
    rng = np.random.default_rng(7)
    v = rng.choice(np.array([-1.0, 1.0, 0.25]), size=A.nnz)
    frac = rng.random(A.nnz) < 0.2
    v[frac] = rng.standard_normal(int(frac.sum()))
    A64 = Csr(A.rows, A.cols, A.row_ptr, A.col_idx, v, dtype=np.float64)
    want = oracle.spgemm_legacy(A64, A64)
    assert int(want[0][-1]) < cases.laplacian_squared_nnz(1024)  # some entries cancelled
    same(rvsp.spgemm_legacy(A64, A64), want)


def test_legacy_int8_power_law(gpu):
    A = gen.rmat(14, seed=3)
    rng = np.random.default_rng(1)
    A8 = Csr(A.rows, A.cols, A.row_ptr, A.col_idx,
             rng.integers(-128, 128, size=A.nnz).astype(np.int8), dtype=np.int8)
    same(rvsp.spgemm_legacy(A8, A8), oracle.spgemm_legacy(A8, A8))


def test_legacy_status_codes(gpu):
    f64 = cases.legacy_pair(1, False, rows=6, inner=5, cols=7)
    i8 = cases.legacy_pair(1, True, rows=6, inner=5, cols=7)
    FP64, INT8, INT32 = F.DTYPE_FP64, F.DTYPE_INT8, F.DTYPE_INT32
    A, B = f64
    assert status(A, B, FP64, FP64) == F.SUCCESS
    # FP64 wrapper: dtypes and shapes are INVALID_ARGUMENT.
    assert status(i8[0], B, FP64, FP64) == F.ERROR_INVALID_ARGUMENT
    assert status(A, A, FP64, FP64) == F.ERROR_INVALID_ARGUMENT  # 6x5 times 6x5
    # FP64 kernel: columns out of range and malformed row pointers.
    bad = Csr(A.rows, A.cols, A.row_ptr, np.where(np.arange(A.nnz) == 0, A.cols, A.col_idx),
              A.values, dtype=np.float64)
    assert status(bad, B, FP64, FP64) == F.ERROR_INVALID_ARGUMENT
    # INT8 wrapper: rvsp_csr_validate first (INVALID_CSR), then dtypes
    # (UNSUPPORTED_DTYPE), then shapes.
    a8, b8 = i8
    assert status(a8, b8, INT8, INT32) == F.SUCCESS
    bad8 = Csr(a8.rows, a8.cols, a8.row_ptr, np.where(np.arange(a8.nnz) == 0, -1, a8.col_idx),
               a8.values, dtype=np.int8)
    assert status(bad8, b8, INT8, INT32) == F.ERROR_INVALID_CSR
    assert status(a8, B, INT8, INT32) == F.ERROR_UNSUPPORTED_DTYPE
    assert status(a8, a8, INT8, INT32) == F.ERROR_INVALID_ARGUMENT
    short = a8.struct()
    short.nnz = a8.nnz - 1  # row_ptr[rows] != nnz
    Cm = F.CsrMatrix()
    opts = F.SpgemmOptions(F.BACKEND_SCALAR, INT8, INT32)
    assert L.rvsp_spgemm_csr(C.byref(short), C.byref(b8.struct()), C.byref(Cm),
                             C.byref(opts)) == F.ERROR_INVALID_CSR
    # The reference library agrees on these codes where it was built.
    if oracle.ref_available():
        assert oracle.ref_spgemm_legacy(bad, B)[0] == F.ERROR_INVALID_ARGUMENT
        assert oracle.ref_spgemm_legacy(bad8, b8)[0] == F.ERROR_INVALID_CSR
        assert oracle.ref_spgemm_legacy(a8, a8)[0] == F.ERROR_INVALID_ARGUMENT
    # Unsorted columns: the legacy kernels accept them, this path reports
    # INVALID_CSR (it needs canonical rows).
    multi = np.flatnonzero(np.diff(a8.row_ptr) > 1)
    ci = a8.col_idx.copy()
    p = int(a8.row_ptr[multi[0]])
    ci[p], ci[p + 1] = ci[p + 1], ci[p]
    uns = Csr(a8.rows, a8.cols, a8.row_ptr, ci, a8.values, dtype=np.int8)
    assert status(uns, b8, INT8, INT32) == F.ERROR_INVALID_CSR
    # Other dtype pairs stay unsupported.
    assert status(A, B, FP64, F.DTYPE_FP32) == F.ERROR_UNSUPPORTED_BACKEND


def test_legacy_fp64_without_rows():
    A = Csr(0, 4, [0], [], [], dtype=np.float64)
    B = Csr(4, 3, [0, 1, 1, 2, 2], [0, 2], [1.0, 2.0], dtype=np.float64)
    got = rvsp.spgemm_legacy(A, B)
    assert got.rows == 0 and got.row_ptr.tolist() == [0] and got.nnz == 0
