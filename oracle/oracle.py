"""CPU oracle for the SpGEMM hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker or the CPU baseline. The product
library never loads it.

Two CPU implementations sit behind it:
  liboracle.so          this project's restatement of the reference scalar
                        algorithm (spgemm_oracle.c; cites the reference lines
                        it follows)
  _ref/librvsp_ref.so   the reference's own C sources compiled as they lie
                        (oracle/Makefile `ref`); present only where the
                        reference tree was available at build time

Pinning (see tests/test_oracle.py): both libraries must agree bit for bit
on seeded inputs and with the committed golden digests; the six hand fixtures
of the reference's tests/csr_fixtures.h:42-154; brute-force dense products on
tiny inputs; the closed-form nnz of the 2-D Laplacian squared; and planted
mistakes (FMA, reversed order) must be caught.
"""
import ctypes as C
import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "librvsp_ref.so")

_P = C.c_void_p
_I32 = C.c_int32


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


_olib = None


def olib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        h = C.CDLL(ORACLE_LIB)
        h.oracle_symbolic_count.argtypes = [_I32, _I32, _I32] + [_P] * 7
        h.oracle_symbolic_fill.argtypes = [_I32, _I32, _I32] + [_P] * 6
        h.oracle_numeric_f32.argtypes = [_I32, _I32, _I32] + [_P] * 9
        h.oracle_numeric_f32_variant.argtypes = [_I32, _I32, _I32] + [_P] * 9 + [C.c_int]
        h.oracle_numeric_f64.argtypes = [_I32, _I32, _I32] + [_P] * 10
        h.oracle_numeric_legacy.argtypes = [_I32, _I32, _I32, C.c_int] + [_P] * 9 + [C.c_int]
        for f in ("oracle_symbolic_count", "oracle_symbolic_fill", "oracle_numeric_f32",
                  "oracle_numeric_f32_variant", "oracle_numeric_f64", "oracle_numeric_legacy"):
            getattr(h, f).restype = C.c_int
        _olib = h
    return _olib


class OracleOverflow(RuntimeError):
    pass


def symbolic(A, B, r0=0, r1=None):
    """Block-local (row_ptr, op_counts) of C's rows [r0, r1)."""
    r1 = A.rows if r1 is None else r1
    rp = np.zeros(r1 - r0 + 1, dtype=np.int32)
    ops = np.zeros(r1 - r0, dtype=np.int64)
    tot = C.c_int64()
    st = olib().oracle_symbolic_count(r0, r1, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                                      _ptr(B.row_ptr), _ptr(B.col_idx), _ptr(rp),
                                      _ptr(ops), C.byref(tot))
    if st == -6:
        raise OracleOverflow("nnz(C) exceeds INT32_MAX")
    if st:
        raise MemoryError("oracle allocation failed")
    return rp, ops


def spgemm(A, B, r0=0, r1=None, variant=0):
    """(row_ptr, col_idx, values) of C's rows [r0, r1), reference semantics."""
    r1 = A.rows if r1 is None else r1
    rp, _ = symbolic(A, B, r0, r1)
    nnz = int(rp[-1])
    ci = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float32)
    h = olib()
    args = (r0, r1, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx))
    if h.oracle_symbolic_fill(*args, _ptr(B.row_ptr), _ptr(B.col_idx), _ptr(rp), _ptr(ci)):
        raise MemoryError("oracle allocation failed")
    if h.oracle_numeric_f32_variant(*args, _ptr(A.values), _ptr(B.row_ptr), _ptr(B.col_idx),
                                    _ptr(B.values), _ptr(rp), _ptr(ci), _ptr(val), variant):
        raise MemoryError("oracle allocation failed")
    return rp, ci[:nnz], val[:nnz]


def spgemm_legacy(A, B, variant=0):
    """(row_ptr, col_idx, values) of the legacy-dtype product, reference
    semantics (src/kernels/spgemm/legacy/csr_scalar_f64.c and
    csr_scalar_i8.c): A, B with float64 values give float64 C, int8 values
    give int32 C; columns ascending, exact-zero sums dropped."""
    i8 = A.values.dtype == np.int8
    rp, _ = symbolic(A, B)
    nnz = int(rp[-1])
    ci = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.int32 if i8 else np.float64)
    h = olib()
    if h.oracle_symbolic_fill(0, A.rows, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                              _ptr(B.row_ptr), _ptr(B.col_idx), _ptr(rp), _ptr(ci)):
        raise MemoryError("oracle allocation failed")
    if h.oracle_numeric_legacy(0, A.rows, B.cols, int(i8), _ptr(A.row_ptr), _ptr(A.col_idx),
                               _ptr(A.values), _ptr(B.row_ptr), _ptr(B.col_idx),
                               _ptr(B.values), _ptr(rp), _ptr(ci), _ptr(val), variant):
        raise MemoryError("oracle allocation failed")
    ci, val = ci[:nnz], val[:nnz]
    keep = val != 0
    row_of = np.repeat(np.arange(A.rows), np.diff(rp))
    out_rp = np.zeros(A.rows + 1, dtype=np.int64)
    out_rp[1:] = np.cumsum(np.bincount(row_of[keep], minlength=A.rows))
    return out_rp.astype(np.int32), ci[keep], val[keep]


def ref_spgemm_legacy(A, B):
    """The reference library's own rvsp_spgemm_csr_scalar_f64 / _i8."""
    h = rlib()
    i8 = A.values.dtype == np.int8
    fn = h.rvsp_spgemm_csr_scalar_i8 if i8 else h.rvsp_spgemm_csr_scalar_f64
    fn.argtypes = [_P, _P, _P]
    fn.restype = C.c_int
    from rvsp import _ffi as F
    a, b, c = A.struct(), B.struct(), F.CsrMatrix()
    st = fn(C.byref(a), C.byref(b), C.byref(c))
    if st:
        return st, None
    try:
        n = c.nnz
        vt = C.c_int32 if i8 else C.c_double
        rp = np.ctypeslib.as_array((C.c_int32 * (c.rows + 1)).from_address(c.row_ptr)).copy()
        ci = np.ctypeslib.as_array((C.c_int32 * max(n, 1)).from_address(c.col_idx))[:n].copy()
        val = np.ctypeslib.as_array((vt * max(n, 1)).from_address(c.values))[:n].copy()
    finally:
        _libc.free(C.c_void_p(c.row_ptr))
        _libc.free(C.c_void_p(c.col_idx))
        _libc.free(C.c_void_p(c.values))
    return 0, (rp, ci, val)


def spgemm_f64(A, B, rp, ci, r0=0):
    """FP64 values and sum_k |a_ik b_kj| for the structure (rp, ci) of rows
    [r0, r0 + len(rp) - 1)."""
    r1 = r0 + rp.size - 1
    n = max(int(rp[-1]), 1)
    v = np.zeros(n, dtype=np.float64)
    mag = np.zeros(n, dtype=np.float64)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    if olib().oracle_numeric_f64(r0, r1, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                                 _ptr(A.values), _ptr(B.row_ptr), _ptr(B.col_idx),
                                 _ptr(B.values), _ptr(rp), _ptr(ci), _ptr(v), _ptr(mag)):
        raise MemoryError("oracle allocation failed")
    return v[:int(rp[-1])], mag[:int(rp[-1])]


def f64_error(A, B, rp, ci, got, r0=0):
    """Error of FP32 values `got` on the structure (rp, ci) of rows
    [r0, ...) against this FP64 oracle: the largest |c - c64| / |c64| over
    entries whose |c64| is at least 1e-6 of their sum of |a*b| (an entry
    cancelled that far has no relative scale; such entries are counted), and
    the condition-aware max |c - c64| / sum_k |a_ik b_kj| over all entries.
    The reference's own validation (bench/bench.c:300-304) compares FP32
    results with 1e-5 absolute + 1e-4 relative slack; the FP64 figures here
    are reported against a 1e-5 tolerance."""
    v64, mag = spgemm_f64(A, B, np.asarray(rp, dtype=np.int32), ci, r0)
    d = np.abs(np.asarray(got, dtype=np.float64) - v64)
    scaled = np.abs(v64) >= 1e-6 * mag
    rel = float((d[scaled] / np.abs(v64[scaled])).max()) if scaled.any() else 0.0
    cond = float((d / np.where(mag > 0, mag, 1.0)).max()) if d.size else 0.0
    return {"entries": int(d.size), "max_rel_err": rel, "max_cond_err": cond,
            "cancelled_entries": int((~scaled).sum())}


def merge_f64_errors(errs, tolerance=1e-5):
    """Combines f64_error results of several samples (None if all empty)."""
    errs = [e for e in errs if e and e["entries"]]
    if not errs:
        return None
    return {"entries": sum(e["entries"] for e in errs),
            "max_rel_err": max(e["max_rel_err"] for e in errs),
            "max_cond_err": max(e["max_cond_err"] for e in errs),
            "cancelled_entries": sum(e["cancelled_entries"] for e in errs),
            "tolerance": tolerance}


# ---------------------------------------------------------------------------
# The reference library built from its own sources.

_rlib = None


REF_MTX_LIB = os.path.join(os.path.dirname(REF_LIB), "libref_mtx.so")


def ref_mtx_available():
    return os.path.exists(REF_MTX_LIB)


class _Vec(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t), ("capacity", C.c_size_t),
                ("elem_size", C.c_size_t)]


class _RefCsr(C.Structure):
    _fields_ = [("row_ptr", C.POINTER(_Vec)), ("col_ind", C.POINTER(_Vec)),
                ("val", C.POINTER(_Vec))]


def ref_read_mtx(path):
    """The reference's assemble_csr_matrix (tools/src/mtx_to_csr_formatter.c
    :153-325) on a file: (row_ptr, col_idx, values) arrays, or None when it
    returns no matrix."""
    h = C.CDLL(REF_MTX_LIB, mode=C.RTLD_LOCAL)
    h.assemble_csr_matrix.restype = _RefCsr
    h.assemble_csr_matrix.argtypes = [C.c_char_p]
    h.vector_free.argtypes = [C.POINTER(_Vec)]
    m = h.assemble_csr_matrix(path.encode())
    if not m.row_ptr:
        return None
    out = []
    for v, dt in ((m.row_ptr, np.int32), (m.col_ind, np.int32), (m.val, np.float32)):
        n = v.contents.size
        out.append(np.ctypeslib.as_array((C.c_uint8 * max(n * 4, 1)).from_address(v.contents.data))
                   [:n * 4].view(dt).copy() if n else np.zeros(0, dt))
        h.vector_free(v)
    return tuple(out)


def _strtof(text):
    """Decimal text to FP32 as the C library's strtof (one rounding; a
    parse through a double would round twice)."""
    libc = C.CDLL(None)
    libc.strtof.restype = C.c_float
    libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
    return np.float32(libc.strtof(text.encode(), None))


def read_mtx(path):
    """This oracle's restatement of the reference converter in plain
    Python/numpy (small files): (rows, cols, row_ptr, col_idx, values).
    Banner pattern/symmetric, 1-based coordinates, the first `L` entries in
    file order, mirrors of off-diagonal entries appended for symmetric
    storage, rows in that order, then per row a stable sort by column and
    repeated columns summed in order in FP32 (mtx_to_csr_formatter.c
    :74-151, 153-325)."""
    with open(path) as f:
        lines = f.read().split("\n")
    words = lines[0].split()
    pattern, symmetric = words[3] == "pattern", words[4] == "symmetric"
    k = 1
    while lines[k].startswith("%"):
        k += 1
    M, N, L = (int(x) for x in lines[k].split())
    toks = " ".join(lines[k + 1:]).split()
    per = 2 if pattern else 3
    ent = []
    for e in range(L):
        t = toks[e * per:(e + 1) * per]
        v = np.float32(1.0) if pattern else _strtof(t[2])
        ent.append((int(t[0]) - 1, int(t[1]) - 1, v))
    if symmetric:
        ent += [(c, r, v) for r, c, v in ent if r != c]
    rows = [[] for _ in range(M)]
    for r, c, v in ent:
        rows[r].append((c, v))
    rp, ci, vals = [0], [], []
    for row in rows:
        row = sorted(row, key=lambda cv: cv[0])  # stable
        i = 0
        while i < len(row):
            c, acc = row[i][0], np.float32(row[i][1])
            i += 1
            while i < len(row) and row[i][0] == c:
                acc = np.float32(acc + np.float32(row[i][1]))
                i += 1
            ci.append(c)
            vals.append(acc)
        rp.append(len(ci))
    return (M, N, np.array(rp, np.int32), np.array(ci, np.int32),
            np.array(vals, np.float32))


def ref_available():
    return os.path.exists(REF_LIB)


def rlib():
    global _rlib
    if _rlib is None:
        h = C.CDLL(REF_LIB)
        h.rvsp_spgemm_scalar_f32.argtypes = [_I32, _I32, _I32] + [_P] * 7 + [C.c_size_t] + \
            [C.POINTER(_P)] * 3 + [C.POINTER(_I32), _P]
        h.rvsp_spgemm_scalar_f32.restype = C.c_int
        h.rvsp_spgemm_buffer_size.argtypes = [_I32, C.POINTER(C.c_size_t)]
        h.rvsp_spgemm_buffer_size.restype = C.c_int
        _rlib = h
    return _rlib


# Canonical-CSR verdicts (values of rvsp_csr_status_t).
CSR_OK, CSR_BAD_ROW_PTR, CSR_COL_OUT_OF_RANGE, CSR_UNSORTED, CSR_DUPLICATE = range(5)


def csr_check(rows, cols, rp, ci, variant=0):
    """(verdict, bad_row) of a CSR structure, restating the reference's
    rvsp_csr_check (src/kernels/spgemm/rvsp_support.c:36-126): row_ptr must
    start at 0; rows are scanned in order, a row whose end is below its start
    is BAD_ROW_PTR at that row, a non-empty row with no column array is
    BAD_ROW_PTR with no row named, and within a row each column must lie in
    [0, cols) (COL_OUT_OF_RANGE), differ from the previous one (DUPLICATE) and
    exceed it (UNSORTED); the first failure wins. rp / ci: arrays or None.
    variant != 0 plants a mistake for the tests that pin this function:
    1 lets a repeated column pass, 2 reports the last failing entry."""
    if rows < 0 or cols < 0 or rp is None:
        return CSR_BAD_ROW_PTR, -1
    rp = np.asarray(rp, np.int64)
    if rp[0] != 0:
        return CSR_BAD_ROW_PTR, -1
    s, e = rp[:rows], rp[1:rows + 1]
    dec = np.flatnonzero(e < s)
    limit = int(dec[0]) if dec.size else rows  # rows before it are monotone
    tail = (CSR_BAD_ROW_PTR, limit) if limit < rows else (CSR_OK, -1)
    lens = e[:limit] - s[:limit]
    if ci is None:
        return (CSR_BAD_ROW_PTR, -1) if np.any(lens > 0) else tail
    n = int(rp[limit]) if limit > 0 else 0
    if n == 0:
        return tail
    col = np.asarray(ci[:n], np.int64)
    row_of = np.repeat(np.arange(limit), lens)
    prev = np.empty(n, np.int64)
    prev[0] = -1
    prev[1:] = col[:-1]
    prev[s[:limit][lens > 0]] = -1  # the first entry of a row has no predecessor
    out = (col < 0) | (col >= cols)
    bad = np.flatnonzero(out | ((col < prev) if variant == 1 else (col <= prev)))
    if bad.size == 0:
        return tail
    p = int(bad[-1] if variant == 2 else bad[0])
    kind = CSR_COL_OUT_OF_RANGE if out[p] else (CSR_DUPLICATE if col[p] == prev[p] else CSR_UNSORTED)
    return kind, int(row_of[p])


def ref_csr_check(rows, cols, rp, ci):
    """The reference library's own rvsp_csr_check on host arrays."""
    h = rlib()
    h.rvsp_csr_check.argtypes = [_I32, _I32, _P, _P, C.POINTER(_I32)]
    h.rvsp_csr_check.restype = C.c_int
    rpa = None if rp is None else np.ascontiguousarray(rp, np.int32)
    cia = None if ci is None else np.ascontiguousarray(ci if len(ci) else [0], np.int32)
    bad = _I32(-7)
    st = h.rvsp_csr_check(rows, cols, None if rpa is None else rpa.ctypes.data,
                          None if cia is None else cia.ctypes.data, C.byref(bad))
    return st, bad.value


_libc = C.CDLL(None)
_libc.free.argtypes = [_P]


def ref_spgemm(A, B, want_ops=False):
    """The reference's raw scalar kernel (gustavson_core_f32.inc:66-194)."""
    h = rlib()
    ws_bytes = C.c_size_t()
    h.rvsp_spgemm_buffer_size(B.cols, C.byref(ws_bytes))
    ws = np.zeros(ws_bytes.value, dtype=np.uint8)
    rp, ci, val = _P(), _P(), _P()
    nnz = _I32()
    ops = np.zeros(A.rows, dtype=np.int64) if want_ops else None
    st = h.rvsp_spgemm_scalar_f32(A.rows, A.cols, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                                  _ptr(A.values), _ptr(B.row_ptr), _ptr(B.col_idx),
                                  _ptr(B.values), _ptr(ws), ws_bytes.value, C.byref(rp),
                                  C.byref(ci), C.byref(val), C.byref(nnz), _ptr(ops))
    if st != 0:
        raise RuntimeError(f"reference kernel returned {st}")
    n = nnz.value
    out_rp = np.ctypeslib.as_array((C.c_int32 * (A.rows + 1)).from_address(rp.value)).copy()
    out_ci = np.ctypeslib.as_array((C.c_int32 * max(n, 1)).from_address(ci.value))[:n].copy()
    out_v = np.ctypeslib.as_array((C.c_float * max(n, 1)).from_address(val.value))[:n].copy()
    for p in (rp, ci, val):
        _libc.free(p)
    return (out_rp, out_ci, out_v, ops) if want_ops else (out_rp, out_ci, out_v)


# ---------------------------------------------------------------------------

def digest(rp, ci, val):
    """SHA-256 of the little-endian bytes of row_ptr, col_idx and values."""
    h = hashlib.sha256()
    for a, dt in ((rp, "<i4"), (ci, "<i4"), (val, "<f4")):
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def dense_product_structure(A, B):
    """Brute force for tiny inputs: pattern of |A|*|B| with ones as values
    (no cancellation possible), as (row_ptr, col_idx)."""
    da = np.zeros((A.rows, A.cols))
    db = np.zeros((B.rows, B.cols))
    ra = np.repeat(np.arange(A.rows), np.diff(A.row_ptr))
    rb = np.repeat(np.arange(B.rows), np.diff(B.row_ptr))
    da[ra, A.col_idx] = 1.0
    db[rb, B.col_idx] = 1.0
    mask = (da @ db) > 0
    rp = np.zeros(A.rows + 1, dtype=np.int32)
    rp[1:] = np.cumsum(mask.sum(axis=1))
    return rp, np.nonzero(mask)[1].astype(np.int32)


# ---------------------------------------------------------------------------
# CPU baselines for bench.py: numeric-only multiplies through a descriptor.

def _ref_descr_api():
    h = rlib()
    if getattr(h, "_descr_ready", False):
        return h
    P = C.c_void_p
    h.rvsp_spgemm_descr_create.argtypes = [C.POINTER(P)]
    h.rvsp_spgemm_descr_destroy.argtypes = [P]
    h.rvsp_spgemm_work_estimation.argtypes = [P, P, P, C.POINTER(C.c_size_t), C.POINTER(_I32)]
    h.rvsp_spgemm_compute.argtypes = [P, P, P, P, P]
    for f in ("rvsp_spgemm_descr_create", "rvsp_spgemm_work_estimation", "rvsp_spgemm_compute"):
        getattr(h, f).restype = C.c_int
    h._descr_ready = True
    return h


def time_reference_numeric(A, B, reps, struct_cls, values_list=None):
    """Seconds per numeric-only compute of the reference library's
    descriptor path (src/core/spgemm_descr.c compute with the fill skipped),
    one thread. Returns (times, nnz_c)."""
    import time
    h = _ref_descr_api()
    d = C.c_void_p()
    h.rvsp_spgemm_descr_create(C.byref(d))
    a = struct_cls(A.rows, A.cols, A.nnz, A.row_ptr.ctypes.data, A.col_idx.ctypes.data,
                   A.values.ctypes.data, 3, 0, 0)
    b = struct_cls(B.rows, B.cols, B.nnz, B.row_ptr.ctypes.data, B.col_idx.ctypes.data,
                   B.values.ctypes.data, 3, 0, 0)
    ws_bytes, nnz = C.c_size_t(), _I32()
    if h.rvsp_spgemm_work_estimation(d, C.byref(a), C.byref(b), C.byref(ws_bytes), C.byref(nnz)):
        raise RuntimeError("reference work_estimation failed")
    n = max(nnz.value, 1)
    rp = np.zeros(A.rows + 1, np.int32)
    ci = np.zeros(n, np.int32)
    v = np.zeros(n, np.float32)
    ws = np.zeros(ws_bytes.value, np.uint8)
    cm = struct_cls(A.rows, B.cols, nnz.value, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                    3, 0, 0)
    h.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(cm), _ptr(ws))  # fill
    times = []
    for r in range(reps):
        if values_list is not None:
            a.values = values_list[r % len(values_list)].ctypes.data
            if A is B:
                b.values = a.values
        t0 = time.perf_counter()
        h.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(cm), _ptr(ws))
        times.append(time.perf_counter() - t0)
    h.rvsp_spgemm_descr_destroy(d)
    return times, nnz.value


def time_reference_numeric_threads(A, B, threads, seconds, struct_cls):
    """All-cores figure of the reference's numeric-only compute: A's rows in
    `threads` contiguous blocks balanced by products, one reference
    descriptor per block (descriptors are independent; the library keeps no
    global state), each block's computes repeated on its own host thread for
    about `seconds`. Returns (products per second, threads)."""
    import threading
    import time
    h = _ref_descr_api()
    lens = np.diff(B.row_ptr).astype(np.int64)
    cs = np.concatenate([[0], np.cumsum(lens[A.col_idx])])
    cum = np.concatenate([[0], np.cumsum(cs[A.row_ptr[1:]] - cs[A.row_ptr[:-1]])])
    cuts = [int(np.searchsorted(cum, cum[-1] * t / threads)) for t in range(threads)] + [A.rows]
    b = struct_cls(B.rows, B.cols, B.nnz, B.row_ptr.ctypes.data, B.col_idx.ctypes.data,
                   B.values.ctypes.data, 3, 0, 0)
    work = []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        if r1 <= r0:
            continue
        o0, o1 = int(A.row_ptr[r0]), int(A.row_ptr[r1])
        rp = np.ascontiguousarray(A.row_ptr[r0:r1 + 1] - o0, dtype=np.int32)
        ci = np.ascontiguousarray(A.col_idx[o0:o1])
        va = np.ascontiguousarray(A.values[o0:o1])
        a = struct_cls(r1 - r0, A.cols, o1 - o0, rp.ctypes.data, ci.ctypes.data, va.ctypes.data,
                       3, 0, 0)
        d = C.c_void_p()
        h.rvsp_spgemm_descr_create(C.byref(d))
        ws_bytes, nnz = C.c_size_t(), _I32()
        if h.rvsp_spgemm_work_estimation(d, C.byref(a), C.byref(b), C.byref(ws_bytes),
                                         C.byref(nnz)):
            raise RuntimeError("reference work_estimation failed")
        n = max(nnz.value, 1)
        out = (np.zeros(r1 - r0 + 1, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32))
        cm = struct_cls(r1 - r0, B.cols, nnz.value, out[0].ctypes.data, out[1].ctypes.data,
                        out[2].ctypes.data, 3, 0, 0)
        ws = np.zeros(ws_bytes.value, np.uint8)
        h.rvsp_spgemm_compute(d, C.byref(a), C.byref(b), C.byref(cm), _ptr(ws))  # fill
        work.append({"d": d, "a": a, "cm": cm, "ws": ws, "keep": (rp, ci, va, out),
                     "ops": int(cum[r1] - cum[r0]), "reps": 0})
    start = threading.Barrier(len(work) + 1)
    stop_at = [0.0]

    def run(w):
        start.wait()
        while time.perf_counter() < stop_at[0] or w["reps"] == 0:
            h.rvsp_spgemm_compute(w["d"], C.byref(w["a"]), C.byref(b), C.byref(w["cm"]),
                                  _ptr(w["ws"]))
            w["reps"] += 1
        w["end"] = time.perf_counter()

    ts = [threading.Thread(target=run, args=(w,)) for w in work]
    for t in ts:
        t.start()
    t0 = time.perf_counter()
    stop_at[0] = t0 + seconds
    start.wait()
    for t in ts:
        t.join()
    wall = max(w["end"] for w in work) - t0
    ops = sum(w["ops"] * w["reps"] for w in work)
    for w in work:
        h.rvsp_spgemm_descr_destroy(w["d"])
    return ops / wall, len(work)


def time_port_numeric(A, B, reps):
    """Seconds per numeric pass of this project's oracle restatement."""
    import time
    rp, _ = symbolic(A, B)
    ci = np.zeros(max(int(rp[-1]), 1), np.int32)
    olib().oracle_symbolic_fill(0, A.rows, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                                _ptr(B.row_ptr), _ptr(B.col_idx), _ptr(rp), _ptr(ci))
    v = np.zeros_like(ci, dtype=np.float32)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        olib().oracle_numeric_f32(0, A.rows, B.cols, _ptr(A.row_ptr), _ptr(A.col_idx),
                                  _ptr(A.values), _ptr(B.row_ptr), _ptr(B.col_idx),
                                  _ptr(B.values), _ptr(rp), _ptr(ci), _ptr(v))
        times.append(time.perf_counter() - t0)
    return times, int(rp[-1])
