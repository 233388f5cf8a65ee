"""The C-ABI library loads, exports exactly what include/*.h declares, and its
host-side logic (argument checks, ordering rules, CSR validation, status
strings) behaves as the reference's. None of these tests launches a kernel;
they run with or without a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import rvsp
from rvsp import _ffi as F
from rvsp.api import Csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
L = F.lib()


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rvsp_\w+)\s*\(", text))
    return names


def test_every_declared_function_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = declared_functions()
    assert declared, "no declarations parsed"
    assert declared - exported == set()
    # Nothing beyond the API leaks out.
    assert {e for e in exported if not e.startswith("rvsp_")} == set()
    assert exported - declared == set()


def test_binding_covers_declarations():
    assert set(F.PROTOTYPES) == declared_functions()


def test_reference_signatures_present():
    # The 13 entry points of the reference's include/rv_sparse.h:27-142.
    ref = {"rvsp_get_version", "rvsp_status_to_string", "rvsp_csr_create",
           "rvsp_csr_validate", "rvsp_csr_destroy", "rvsp_spgemm_csr",
           "rvsp_spgemm_descr_create", "rvsp_spgemm_descr_destroy",
           "rvsp_spgemm_set_algo", "rvsp_spgemm_work_estimation",
           "rvsp_spgemm_compute", "rvsp_spgemm_get_op_counts",
           "rvsp_spgemm_invalidate_structure"}
    assert ref <= declared_functions()


def test_struct_layout_matches_reference_abi():
    # rvsp_csr_matrix_t on LP64: 3 int32, 3 pointers, 2 enums, 1 int.
    assert C.sizeof(F.CsrMatrix) == 56
    assert F.CsrMatrix.row_ptr.offset == 16 and F.CsrMatrix.dtype.offset == 40
    assert C.sizeof(F.SpgemmOptions) == 12


def test_version_and_status_strings():
    assert L.rvsp_get_version() == b"0.1.0"
    want = {0: b"success", -1: b"null pointer", -2: b"invalid argument",
            -3: b"invalid CSR matrix", -4: b"unsupported data type",
            -5: b"unsupported backend", -6: b"allocation failed",
            -7: b"device error", 5: b"unknown error"}
    for code, s in want.items():
        assert L.rvsp_status_to_string(code) == s
    assert L.rvsp_csr_status_string(F.CSR_OK) == b"canonical CSR"
    assert L.rvsp_csr_status_string(F.CSR_DUPLICATE) == b"duplicate column in row"
    assert L.rvsp_csr_status_string(99) == b"unknown"


def _mat(rows, cols, rp, ci, vals):
    rp = np.asarray(rp, np.int32)
    ci = np.asarray(ci, np.int32)
    v = np.asarray(vals, np.float32)
    m = F.CsrMatrix()
    st = L.rvsp_csr_create(C.byref(m), rows, cols, len(ci), rp.ctypes.data,
                           ci.ctypes.data, v.ctypes.data, F.DTYPE_FP32)
    return st, m, (rp, ci, v)


def test_csr_create_and_validate():
    st, m, keep = _mat(2, 2, [0, 2, 3], [0, 1, 1], [1, 2, 3])
    assert st == F.SUCCESS and m.owns_data == 0 and m.nnz == 3
    st, m, keep = _mat(2, 2, [0, 2, 3], [0, 1, 2], [1, 2, 3])  # column out of range
    assert st == F.ERROR_INVALID_CSR
    st, m, keep = _mat(2, 2, [1, 2, 3], [0, 1, 1], [1, 2, 3])  # row_ptr[0] != 0
    assert st == F.ERROR_INVALID_CSR
    st, m, keep = _mat(2, 2, [0, 2, 3], [1, 0, 1], [1, 2, 3])  # unsorted: not checked here
    assert st == F.SUCCESS
    m = F.CsrMatrix()
    assert L.rvsp_csr_create(None, 1, 1, 0, 1, 1, 1, F.DTYPE_FP32) == F.ERROR_NULL_POINTER
    rp = np.zeros(2, np.int32)
    assert L.rvsp_csr_create(C.byref(m), 0, 1, 0, rp.ctypes.data, rp.ctypes.data,
                             rp.ctypes.data, F.DTYPE_FP32) == F.ERROR_INVALID_ARGUMENT
    L.rvsp_csr_destroy(None)
    L.rvsp_csr_destroy(C.byref(m))
    assert m.rows == 0 and m.row_ptr is None


def test_csr_check_verdicts():
    def chk(rows, cols, rp, ci):
        rp = np.asarray(rp, np.int32)
        ci = np.asarray(ci if len(ci) else [0], np.int32)
        bad = C.c_int32(7)
        st = L.rvsp_csr_check(rows, cols, rp.ctypes.data, ci.ctypes.data, C.byref(bad))
        return st, bad.value
    assert chk(2, 3, [0, 2, 3], [0, 2, 1]) == (F.CSR_OK, -1)
    assert chk(2, 3, [0, 2, 1], [0, 2, 1]) == (F.CSR_BAD_ROW_PTR, 1)
    assert chk(2, 3, [0, 2, 3], [0, 3, 1]) == (F.CSR_COL_OUT_OF_RANGE, 0)
    assert chk(2, 3, [0, 2, 3], [1, 1, 1]) == (F.CSR_DUPLICATE, 0)
    assert chk(2, 3, [0, 2, 3], [2, 1, 1]) == (F.CSR_UNSORTED, 0)
    assert chk(1, 3, [1, 1], [])[0] == F.CSR_BAD_ROW_PTR
    assert L.rvsp_csr_check(-1, 3, None, None, None) == F.CSR_BAD_ROW_PTR


def test_buffer_size_matches_reference_layout():
    n = C.c_size_t()
    assert L.rvsp_spgemm_buffer_size(1000, C.byref(n)) == F.SUCCESS
    up = lambda x: (x + 63) // 64 * 64  # noqa: E731
    assert n.value == up(4000) + up(1000) + up(4000) + up(4000)
    assert L.rvsp_spgemm_buffer_size(0, C.byref(n)) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_buffer_size(10, None) == F.ERROR_NULL_POINTER


def test_descriptor_contract_without_device():
    # Reference tests/test_spgemm_descr_f32.c:65-100: creation, algo
    # selection and ordering errors, all decided before any device work.
    d = C.c_void_p()
    assert L.rvsp_spgemm_descr_create(None) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_descr_create(C.byref(d)) == F.SUCCESS and d.value
    assert L.rvsp_spgemm_set_algo(d, F.ALGO_DEFAULT) == F.SUCCESS
    assert L.rvsp_spgemm_set_algo(None, F.ALGO_DEFAULT) == F.ERROR_NULL_POINTER
    for algo in (F.ALGO_RVV, F.ALGO_CONTIG, F.ALGO_ADAPTIVE, 99):
        assert L.rvsp_spgemm_set_algo(d, algo) == F.ERROR_UNSUPPORTED_BACKEND
    for algo in (F.ALGO_HIP_FIXED, F.ALGO_HIP_TUNED, F.ALGO_HIP_LEARNED, F.ALGO_DEFAULT):
        assert L.rvsp_spgemm_set_algo(d, algo) == F.SUCCESS
    A = Csr(2, 2, [0, 2, 3], [0, 1, 1], [1, 2, 3])
    B = Csr(2, 2, [0, 1, 2], [0, 1], [4, 5])
    a, b = A.struct(), B.struct()
    rp, ci, v = np.zeros(3, np.int32), np.zeros(3, np.int32), np.zeros(3, np.float32)
    Cm = F.CsrMatrix(2, 2, 3, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                     F.DTYPE_FP32, F.FORMAT_CSR, 0)
    ws = np.zeros(4096, np.uint8)
    assert L.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(Cm),
                                 ws.ctypes.data) == F.ERROR_INVALID_ARGUMENT
    ops = np.zeros(2, np.int64)
    assert L.rvsp_spgemm_get_op_counts(d, ops.ctypes.data, 2) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_compute(None, C.byref(a), C.byref(b), C.byref(Cm),
                                 ws.ctypes.data) == F.ERROR_NULL_POINTER
    sz, nnz = C.c_size_t(), C.c_int32()
    assert L.rvsp_spgemm_work_estimation(d, None, C.byref(b), C.byref(sz),
                                         C.byref(nnz)) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_work_estimation(d, C.byref(a), C.byref(b), None,
                                         C.byref(nnz)) == F.ERROR_NULL_POINTER
    bad = A.struct()
    bad.dtype = F.DTYPE_FP64
    assert L.rvsp_spgemm_work_estimation(d, C.byref(bad), C.byref(b), C.byref(sz),
                                         C.byref(nnz)) == F.ERROR_UNSUPPORTED_DTYPE
    wrong = Csr(3, 2, [0, 1, 2, 2], [0, 1], [1, 1]).struct()
    assert L.rvsp_spgemm_work_estimation(d, C.byref(a), C.byref(wrong), C.byref(sz),
                                         C.byref(nnz)) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_set_memory(d, 5) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_set_memory(d, F.MEMORY_DEVICE) == F.SUCCESS
    L.rvsp_spgemm_invalidate_structure(None)
    L.rvsp_spgemm_descr_destroy(None)
    L.rvsp_spgemm_descr_destroy(d)


def test_oneshot_rejects_before_device_work():
    A = Csr(2, 2, [0, 2, 3], [1, 0, 1], [1, 2, 3])  # unsorted row: not canonical
    B = Csr(2, 2, [0, 1, 2], [0, 1], [4, 5])
    a, b, Cm = A.struct(), B.struct(), F.CsrMatrix()
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(Cm), None) == F.ERROR_INVALID_CSR
    # FP64 options reach the FP64 wrapper, which refuses FP32 operands
    # (csr_spgemm_wrappers.c:362-365) before any device work.
    opts = F.SpgemmOptions(F.BACKEND_SCALAR, F.DTYPE_FP64, F.DTYPE_FP64)
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(Cm),
                             C.byref(opts)) == F.ERROR_INVALID_ARGUMENT
    opts = F.SpgemmOptions(F.BACKEND_SCALAR, F.DTYPE_FP64, F.DTYPE_FP32)
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(Cm),
                             C.byref(opts)) == F.ERROR_UNSUPPORTED_BACKEND
    assert L.rvsp_spgemm_csr(None, C.byref(b), C.byref(Cm), None) == F.ERROR_NULL_POINTER


@pytest.mark.skipif(rvsp.device_count() > 0, reason="a GPU is present")
def test_no_device_fails_loudly():
    # Without a gfx950 device the library reports an error; it has no CPU path.
    A = Csr(2, 2, [0, 2, 3], [0, 1, 1], [1, 2, 3])
    with pytest.raises(rvsp.RvspError) as e:
        rvsp.spgemm(A, A)
    assert e.value.status == F.ERROR_DEVICE


def test_batch_entry_points_check_arguments_without_a_device():
    h = C.c_void_p()
    F.check(L.rvsp_spgemm_descr_create(C.byref(h)))
    a = (F.CsrMatrix * 1)()
    out = (C.c_int32 * 1)()
    assert L.rvsp_spgemm_batch_work_estimation(None, 1, a, a, out) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_batch_work_estimation(h, 1, a, a, None) == F.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_batch_work_estimation(h, 0, a, a, out) == F.ERROR_INVALID_ARGUMENT
    # Host memory (the default) has no batched path.
    assert L.rvsp_spgemm_batch_work_estimation(h, 1, a, a, out) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_batch_compute(h, 1, a, a, a) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_batch_groups(h) == -1
    assert L.rvsp_spgemm_batch_graph_launches(h) == 0
    assert L.rvsp_spgemm_batch_set_group_limit(h, -1) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_batch_set_group_limit(h, 1 << 40) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_spgemm_batch_set_group_limit(h, 1000) == F.SUCCESS
    assert L.rvsp_spgemm_batch_set_graph(h, 0) == F.SUCCESS
    assert L.rvsp_spgemm_batch_set_graph(None, 0) == F.ERROR_NULL_POINTER
    L.rvsp_spgemm_descr_destroy(h)
