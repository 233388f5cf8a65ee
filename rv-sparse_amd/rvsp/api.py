"""Python host mirror of the rv-sparse C API, over librvsparse_hip.so.

The names follow the reference's operator interface (include/rv_sparse.h):
a CSR container, the one-shot multiply, and the two-phase descriptor with
its selectable strategy. Every call goes through the C-ABI library; every
multiply runs on the GPU.
"""
import ctypes as C

import numpy as np

from . import _ffi as F
from ._ffi import check, lib


_DTYPES = {np.dtype(np.float32): F.DTYPE_FP32, np.dtype(np.float64): F.DTYPE_FP64,
           np.dtype(np.int8): F.DTYPE_INT8, np.dtype(np.int32): F.DTYPE_INT32}


class Csr:
    """Host CSR matrix (int32 row_ptr/col_idx; float32 values unless another
    value dtype of the C API is given: float64, int8 or int32)."""

    def __init__(self, rows, cols, row_ptr, col_idx, values, dtype=np.float32):
        self.rows = int(rows)
        self.cols = int(cols)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=dtype)

    @property
    def nnz(self):
        return int(self.col_idx.size)

    def struct(self):
        """rvsp_csr_matrix_t viewing this object's arrays (keep self alive)."""
        return F.CsrMatrix(self.rows, self.cols, self.nnz,
                           self.row_ptr.ctypes.data, self.col_idx.ctypes.data,
                           self.values.ctypes.data, _DTYPES[self.values.dtype],
                           F.FORMAT_CSR, 0)

    @classmethod
    def from_dense(cls, dense):
        dense = np.asarray(dense, dtype=np.float32)
        rows, cols = dense.shape
        mask = dense != 0
        row_ptr = np.zeros(rows + 1, dtype=np.int32)
        row_ptr[1:] = np.cumsum(mask.sum(axis=1))
        r, c = np.nonzero(mask)
        return cls(rows, cols, row_ptr, c.astype(np.int32), dense[r, c])

    def to_dense(self):
        out = np.zeros((self.rows, self.cols), dtype=np.float32)
        r = np.repeat(np.arange(self.rows), np.diff(self.row_ptr))
        out[r, self.col_idx] = self.values
        return out

    def row_ids(self):
        return np.repeat(np.arange(self.rows, dtype=np.int64), np.diff(self.row_ptr))

    def with_values(self, values):
        return Csr(self.rows, self.cols, self.row_ptr, self.col_idx, values)

    def __repr__(self):
        return f"Csr({self.rows}x{self.cols}, nnz={self.nnz})"


# ---------------------------------------------------------------------------
# Device memory through the library (no torch, no separate HIP binding).

def device_alloc(nbytes):
    p = C.c_void_p()
    check(lib().rvsp_hip_malloc(C.byref(p), int(max(nbytes, 4))), "hip_malloc")
    return p.value


def device_free(ptr):
    if ptr:
        check(lib().rvsp_hip_free(C.c_void_p(ptr)), "hip_free")


def to_device(ptr, array):
    array = np.ascontiguousarray(array)
    if array.nbytes:
        check(lib().rvsp_hip_memcpy(C.c_void_p(ptr), C.c_void_p(array.ctypes.data),
                                    array.nbytes, 1), "memcpy h2d")


def from_device(ptr, count, dtype):
    out = np.empty(int(count), dtype=dtype)
    if out.nbytes:
        check(lib().rvsp_hip_memcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr),
                                    out.nbytes, 2), "memcpy d2h")
    return out


def synchronize():
    check(lib().rvsp_hip_synchronize(), "synchronize")


def csr_check_device(rows, cols, nnz, row_ptr, col_idx):
    """rvsp_csr_check_device on device pointers (ints; col_idx may be None):
    returns (verdict, bad_row) with rvsp_csr_check's meaning."""
    st, bad = C.c_int32(-1), C.c_int32(-7)
    check(lib().rvsp_csr_check_device(rows, cols, nnz, row_ptr, col_idx, C.byref(st),
                                      C.byref(bad)), "csr_check_device")
    return st.value, bad.value


def device_count():
    return int(lib().rvsp_hip_device_count())


class DeviceCsr:
    """CSR matrix whose arrays live in device memory."""

    def __init__(self, rows, cols, nnz, alloc_values=True):
        self.rows, self.cols, self.nnz = int(rows), int(cols), int(nnz)
        self.row_ptr = device_alloc(4 * (self.rows + 1))
        self.col_idx = device_alloc(4 * max(self.nnz, 1))
        self.values = device_alloc(4 * max(self.nnz, 1)) if alloc_values else None

    @classmethod
    def upload(cls, host):
        d = cls(host.rows, host.cols, host.nnz)
        to_device(d.row_ptr, host.row_ptr)
        to_device(d.col_idx, host.col_idx)
        to_device(d.values, host.values)
        return d

    def set_values(self, values):
        to_device(self.values, np.ascontiguousarray(values, dtype=np.float32))

    def struct(self, nnz=None):
        return F.CsrMatrix(self.rows, self.cols, self.nnz if nnz is None else nnz,
                           self.row_ptr, self.col_idx, self.values,
                           F.DTYPE_FP32, F.FORMAT_CSR, 0)

    def download(self):
        rp = from_device(self.row_ptr, self.rows + 1, np.int32)
        ci = from_device(self.col_idx, self.nnz, np.int32)
        v = from_device(self.values, self.nnz, np.float32)
        return Csr(self.rows, self.cols, rp, ci, v)

    def free(self):
        for name in ("row_ptr", "col_idx", "values"):
            device_free(getattr(self, name))
            setattr(self, name, None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------

class Descriptor:
    """The two-phase SpGEMM descriptor (reference src/core/spgemm_descr.c)."""

    def __init__(self, algo=F.ALGO_DEFAULT, memory=F.MEMORY_HOST, choices=None):
        self.handle = C.c_void_p()
        check(lib().rvsp_spgemm_descr_create(C.byref(self.handle)), "descr_create")
        check(lib().rvsp_spgemm_set_algo(self.handle, algo), "set_algo")
        check(lib().rvsp_spgemm_set_memory(self.handle, memory), "set_memory")
        if choices:
            arr = (F.BinChoice * len(choices))(*[F.BinChoice(*c) for c in choices])
            check(lib().rvsp_spgemm_set_bin_choices(self.handle, arr, len(choices)),
                  "set_bin_choices")
        self.memory = memory
        self.c_nnz = None
        self.ws = np.zeros(64, dtype=np.uint8)

    def estimate(self, A, B):
        self._a, self._b = A.struct(), B.struct()
        ws = C.c_size_t()
        nnz = C.c_int32()
        check(lib().rvsp_spgemm_work_estimation(self.handle, C.byref(self._a),
                                                C.byref(self._b), C.byref(ws),
                                                C.byref(nnz)), "work_estimation")
        self.ws = np.zeros(max(int(ws.value), 1), dtype=np.uint8)
        self.c_nnz = int(nnz.value)
        return self.c_nnz

    def compute(self, A, B, Cmat):
        """Cmat: rvsp_csr_matrix_t with storage; returns the status."""
        a, b = A.struct(), B.struct()
        return lib().rvsp_spgemm_compute(self.handle, C.byref(a), C.byref(b),
                                         C.byref(Cmat), C.c_void_p(self.ws.ctypes.data))

    def op_counts(self, rows):
        out = np.zeros(rows, dtype=np.int64)
        check(lib().rvsp_spgemm_get_op_counts(self.handle, C.c_void_p(out.ctypes.data),
                                              rows), "get_op_counts")
        return out

    def plan(self):
        n = C.c_int32()
        check(lib().rvsp_spgemm_get_plan(self.handle, None, 0, C.byref(n)), "get_plan")
        arr = (F.BinInfo * n.value)()
        check(lib().rvsp_spgemm_get_plan(self.handle, arr, n.value, C.byref(n)), "get_plan")
        return [dict(phase=b.phase, bin=b.bin, kind=F.KIND_NAMES.get(b.kind, "?"),
                     group=b.group, cap=b.cap, rows=b.rows, lds_bytes=b.lds_bytes,
                     bound=b.bound) for b in arr]

    def stream(self):
        s = C.c_void_p()
        check(lib().rvsp_spgemm_get_stream(self.handle, C.byref(s)), "get_stream")
        return s.value

    def set_async(self, on=True):
        check(lib().rvsp_spgemm_set_async(self.handle, 1 if on else 0), "set_async")

    def invalidate(self):
        lib().rvsp_spgemm_invalidate_structure(self.handle)

    def close(self):
        if self.handle:
            lib().rvsp_spgemm_descr_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spgemm(A, B, algo=F.ALGO_DEFAULT, choices=None):
    """C = A * B through the descriptor path with host operands."""
    d = Descriptor(algo=algo, choices=choices)
    try:
        nnz = d.estimate(A, B)
        n = max(nnz, 1)
        rp = np.zeros(A.rows + 1, dtype=np.int32)
        ci = np.zeros(n, dtype=np.int32)
        v = np.zeros(n, dtype=np.float32)
        Cm = F.CsrMatrix(A.rows, B.cols, nnz, rp.ctypes.data, ci.ctypes.data,
                         v.ctypes.data, F.DTYPE_FP32, F.FORMAT_CSR, 0)
        check(d.compute(A, B, Cm), "compute")
        return Csr(A.rows, B.cols, rp, ci[:nnz], v[:nnz])
    finally:
        d.close()


def spgemm_oneshot(A, B, backend=F.BACKEND_SCALAR):
    """rvsp_spgemm_csr: the library allocates C; copied out, then destroyed."""
    a, b = A.struct(), B.struct()
    Cm = F.CsrMatrix()
    opts = F.SpgemmOptions(backend, F.DTYPE_FP32, F.DTYPE_FP32)
    check(lib().rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(Cm), C.byref(opts)),
          "spgemm_csr")
    try:
        nnz = Cm.nnz
        rp = np.ctypeslib.as_array((C.c_int32 * (Cm.rows + 1)).from_address(Cm.row_ptr)).copy()
        ci = np.ctypeslib.as_array((C.c_int32 * max(nnz, 1)).from_address(Cm.col_idx))[:nnz].copy()
        v = np.ctypeslib.as_array((C.c_float * max(nnz, 1)).from_address(Cm.values))[:nnz].copy()
    finally:
        lib().rvsp_csr_destroy(C.byref(Cm))
    return Csr(A.rows, B.cols, rp, ci, v)


def spgemm_legacy(A, B, backend=F.BACKEND_SCALAR):
    """rvsp_spgemm_csr for float64 operands (FP64 -> FP64) or int8 operands
    (INT8 -> INT32); returns C with float64 / int32 values."""
    a, b = A.struct(), B.struct()
    i8 = A.values.dtype == np.int8
    dt_in, dt_out = (F.DTYPE_INT8, F.DTYPE_INT32) if i8 else (F.DTYPE_FP64, F.DTYPE_FP64)
    Cm = F.CsrMatrix()
    opts = F.SpgemmOptions(backend, dt_in, dt_out)
    check(lib().rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(Cm), C.byref(opts)),
          "spgemm_csr")
    try:
        nnz = Cm.nnz
        vt = C.c_int32 if i8 else C.c_double
        rp = np.ctypeslib.as_array((C.c_int32 * (Cm.rows + 1)).from_address(Cm.row_ptr)).copy()
        ci = np.ctypeslib.as_array((C.c_int32 * max(nnz, 1)).from_address(Cm.col_idx))[:nnz].copy()
        v = np.ctypeslib.as_array((vt * max(nnz, 1)).from_address(Cm.values))[:nnz].copy()
        assert Cm.dtype == dt_out
    finally:
        lib().rvsp_csr_destroy(C.byref(Cm))
    return Csr(A.rows, B.cols, rp, ci, v, dtype=np.int32 if i8 else np.float64)


def spgemm_device(A, B, algo=F.ALGO_DEFAULT, choices=None, repeats=1):
    """C = A * B with device-resident operands; returns (C, descriptor plan)."""
    dA = DeviceCsr.upload(A)
    dB = dA if B is A else DeviceCsr.upload(B)
    d = Descriptor(algo=algo, memory=F.MEMORY_DEVICE, choices=choices)
    try:
        nnz = d.estimate(dA, dB)
        dC = DeviceCsr(A.rows, B.cols, nnz)
        try:
            Cm = dC.struct(nnz)
            for _ in range(repeats):
                check(d.compute(dA, dB, Cm), "compute")
            return dC.download(), d.plan()
        finally:
            dC.free()
    finally:
        d.close()
        dA.free()
        if dB is not dA:
            dB.free()
