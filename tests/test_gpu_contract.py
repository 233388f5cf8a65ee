"""The reference's descriptor contract test (tests/test_spgemm_descr_f32.c,
27 checks) and strategy sweep (tests/test_spgemm_reference_f32.c) restated
against the HIP library, in both operand-memory modes."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import rvsp
from rvsp import _ffi as F
from rvsp.api import Csr

pytestmark = pytest.mark.gpu
L = F.lib()
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["fixtures"]


def fx_csr(m):
    return Csr(m["rows"], m["cols"], m["row_ptr"], m["col_idx"], m["values"])


def test_descriptor_contract(gpu):
    fx = FIXTURES[0]  # "basic"
    A, B = fx_csr(fx["a"]), fx_csr(fx["b"])
    a, b = A.struct(), B.struct()
    d = C.c_void_p()
    assert L.rvsp_spgemm_descr_create(C.byref(d)) == F.SUCCESS and d.value
    assert L.rvsp_spgemm_set_algo(d, F.ALGO_DEFAULT) == F.SUCCESS

    ws_bytes, c_nnz = C.c_size_t(), C.c_int32()
    assert L.rvsp_spgemm_work_estimation(d, C.byref(a), C.byref(b), C.byref(ws_bytes),
                                         C.byref(c_nnz)) == F.SUCCESS
    assert c_nnz.value == 3 and ws_bytes.value > 0
    assert L.rvsp_spgemm_work_estimation(d, None, C.byref(b), C.byref(ws_bytes),
                                         C.byref(c_nnz)) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_set_algo(d, F.ALGO_DEFAULT) == F.ERROR_INVALID_ARGUMENT
    ops = np.zeros(2, np.int64)
    assert L.rvsp_spgemm_get_op_counts(d, ops.ctypes.data, 2) == F.SUCCESS
    assert ops.tolist() == [2, 1]
    assert L.rvsp_spgemm_get_op_counts(d, ops.ctypes.data, 1) == F.ERROR_INVALID_ARGUMENT

    n = c_nnz.value
    rp, ci, v = np.zeros(3, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
    ws = np.zeros(ws_bytes.value, np.uint8)
    Cm = F.CsrMatrix(2, 2, n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                     F.DTYPE_FP32, F.FORMAT_CSR, 0)
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(Cm), None) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), None, ws.ctypes.data) == F.ERROR_NULL_POINTER
    small = F.CsrMatrix.from_buffer_copy(Cm)
    small.nnz = n - 1
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(small),
                                 ws.ctypes.data) == F.ERROR_INVALID_ARGUMENT
    wrong = F.CsrMatrix.from_buffer_copy(Cm)
    wrong.rows = 3
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(wrong),
                                 ws.ctypes.data) == F.ERROR_INVALID_ARGUMENT
    bad_a = F.CsrMatrix.from_buffer_copy(a)
    bad_a.nnz = 2
    assert L.rvsp_spgemm_compute(d, C.byref(bad_a), C.byref(b), C.byref(Cm),
                                 ws.ctypes.data) == F.ERROR_INVALID_ARGUMENT
    fp64 = F.CsrMatrix.from_buffer_copy(Cm)
    fp64.dtype = F.DTYPE_FP64
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(fp64),
                                 ws.ctypes.data) == F.ERROR_UNSUPPORTED_DTYPE

    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(Cm), ws.ctypes.data) == F.SUCCESS
    assert Cm.nnz == 3
    assert ci.tolist() == [0, 1, 1] and v.tolist() == [4.0, 10.0, 15.0]
    assert rp.tolist() == [0, 2, 3]

    # Reuse: doubled A values, same descriptor and C, no re-analysis.
    A.values *= 2
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(Cm), ws.ctypes.data) == F.SUCCESS
    assert v.tolist() == [8.0, 20.0, 30.0]
    assert ci.tolist() == [0, 1, 1]

    L.rvsp_spgemm_invalidate_structure(d)
    ci[:] = 0
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(Cm), ws.ctypes.data) == F.SUCCESS
    assert ci.tolist() == [0, 1, 1]
    L.rvsp_spgemm_descr_destroy(d)


def test_reestimation_replaces_structure(gpu):
    A, B = fx_csr(FIXTURES[0]["a"]), fx_csr(FIXTURES[0]["b"])
    A2, B2 = fx_csr(FIXTURES[5]["a"]), fx_csr(FIXTURES[5]["b"])
    d = rvsp.Descriptor()
    assert d.estimate(A, B) == 3
    assert d.estimate(A2, B2) == 3
    rp = np.zeros(2, np.int32)
    ci = np.zeros(3, np.int32)
    v = np.zeros(3, np.float32)
    Cm = F.CsrMatrix(1, 3, 3, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                     F.DTYPE_FP32, F.FORMAT_CSR, 0)
    F.check(d.compute(A2, B2, Cm))
    assert v.tolist() == [4.0, 2.0, 4.0]
    d.close()


@pytest.mark.parametrize("algo", [F.ALGO_DEFAULT, F.ALGO_RVV, F.ALGO_CONTIG,
                                  F.ALGO_ADAPTIVE, F.ALGO_HIP_FIXED, F.ALGO_HIP_TUNED,
                                  F.ALGO_HIP_LEARNED])
def test_strategy_sweep(gpu, algo):
    # Reference tests/test_spgemm_reference_f32.c:166-186: a strategy the
    # build lacks reports UNSUPPORTED_BACKEND at set_algo (skipped, never
    # passed); every other strategy runs all fixtures.
    d = C.c_void_p()
    L.rvsp_spgemm_descr_create(C.byref(d))
    st = L.rvsp_spgemm_set_algo(d, algo)
    L.rvsp_spgemm_descr_destroy(d)
    if st == F.ERROR_UNSUPPORTED_BACKEND:
        assert algo in (F.ALGO_RVV, F.ALGO_CONTIG, F.ALGO_ADAPTIVE)
        pytest.skip("strategy not provided by this build")
    for fx in FIXTURES:
        got = rvsp.spgemm(fx_csr(fx["a"]), fx_csr(fx["b"]), algo=algo)
        want = fx_csr(fx["c"])
        assert got.row_ptr.tolist() == want.row_ptr.tolist()
        assert got.col_idx.tolist() == want.col_idx.tolist()
        assert np.all(np.abs(got.values - want.values) <= 1e-5 + 1e-4 * np.abs(want.values))


def test_device_mode_contract(gpu):
    A, B = fx_csr(FIXTURES[0]["a"]), fx_csr(FIXTURES[0]["b"])
    dA, dB = rvsp.DeviceCsr.upload(A), rvsp.DeviceCsr.upload(B)
    d = rvsp.Descriptor(memory=F.MEMORY_DEVICE)
    nnz = d.estimate(dA, dB)
    assert L.rvsp_spgemm_set_memory(d.handle, F.MEMORY_HOST) == F.ERROR_INVALID_ARGUMENT
    dC = rvsp.DeviceCsr(2, 2, nnz)
    Cm = dC.struct(nnz)
    F.check(d.compute(dA, dB, Cm))
    got = dC.download()
    assert got.row_ptr.tolist() == [0, 2, 3]
    assert got.col_idx.tolist() == [0, 1, 1] and got.values.tolist() == [4.0, 10.0, 15.0]
    d.close()
    for m in (dA, dB, dC):
        m.free()
