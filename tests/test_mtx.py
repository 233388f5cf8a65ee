"""rvsp_csr_read_mtx and rvsp_csr_canonicalize (include/rv_sparse_mtx.h)
against the reference's Matrix Market converter (tools/src/
mtx_to_csr_formatter.c, compiled from its sources into oracle/_ref) and the
oracle's plain-Python restatement of it, on Matrix Market text written here:
general / symmetric / pattern / integer storage, repeated entries, empty
rows, unsorted entries, decimal values that need one correct rounding to
FP32, and a file large enough for the parallel parse. The restatement is
pinned to the reference bit for bit; scipy's reader pins the structure and
(to FP64 rounding) the values independently."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle
from rvsp import _ffi as F

L = F.lib()
ref = pytest.mark.skipif(not oracle.ref_mtx_available(),
                         reason="reference converter not built here (no reference tree)")


def write(path, banner, rows, cols, entries, comments=("% written by the tests",), count=None):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {banner}\n")
        for c in comments:
            f.write(c + "\n")
        f.write(f"{rows} {cols} {len(entries) if count is None else count}\n")
        for e in entries:
            f.write(" ".join(str(x) for x in e) + "\n")
    return str(path)


def load(path, threads=0):
    m = F.CsrMatrix()
    st = L.rvsp_csr_read_mtx(path.encode(), threads, C.byref(m))
    if st != F.SUCCESS:
        return st, None
    rp = np.ctypeslib.as_array((C.c_int32 * (m.rows + 1)).from_address(m.row_ptr)).copy()
    n = m.nnz
    ci = np.ctypeslib.as_array((C.c_int32 * max(n, 1)).from_address(m.col_idx))[:n].copy()
    v = np.ctypeslib.as_array((C.c_float * max(n, 1)).from_address(m.values))[:n].copy()
    shape = (m.rows, m.cols, m.owns_data, m.dtype)
    L.rvsp_csr_destroy(C.byref(m))
    return st, (shape, rp, ci, v)


def random_entries(rng, rows, cols, n, pattern=False, lower=False):
    out = []
    for _ in range(n):
        r = int(rng.integers(1, rows + 1))
        c = int(rng.integers(1, (r if lower else cols) + 1))
        if pattern:
            out.append((r, c))
        else:
            out.append((r, c, repr(float(rng.uniform(-10, 10)))))
    return out


CASES = {
    "general": ("real general", 7, 9, [(3, 4, "1.5"), (1, 9, "-2.25"), (3, 1, "0.1"),
                                       (3, 4, "0.2"), (7, 7, "1e-3"), (3, 4, "0.3"),
                                       (1, 2, "3.4028235e38"), (7, 1, "1e-45")]),
    "empty_rows": ("real general", 6, 3, [(6, 3, "1"), (2, 2, "2"), (6, 1, "-1")]),
    "symmetric": ("real symmetric", 5, 5, [(1, 1, "2"), (3, 1, "0.5"), (5, 2, "-1.25"),
                                           (3, 1, "0.25"), (4, 4, "7"), (5, 3, "0.1")]),
    "pattern": ("pattern general", 4, 6, [(1, 6), (1, 2), (4, 3), (1, 6), (2, 2)]),
    "pattern_symmetric": ("pattern symmetric", 4, 4, [(2, 1), (3, 3), (4, 1), (4, 2)]),
    "integer": ("integer general", 3, 3, [(1, 1, "3"), (2, 3, "-4"), (1, 1, "5")]),
}


def check_against_restatement(path):
    st, got = load(path)
    assert st == F.SUCCESS
    (rows, cols, owns, dtype), rp, ci, v = got
    M, N, wrp, wci, wv = oracle.read_mtx(path)
    assert (rows, cols, owns, dtype) == (M, N, 1, F.DTYPE_FP32)
    assert rp.tobytes() == wrp.tobytes() and ci.tobytes() == wci.tobytes()
    assert v.tobytes() == wv.tobytes()
    return rp, ci, v


@pytest.mark.parametrize("name", sorted(CASES))
def test_small_files_match_restatement(tmp_path, name):
    banner, rows, cols, ent = CASES[name]
    check_against_restatement(write(tmp_path / f"{name}.mtx", banner, rows, cols, ent))


@ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_small_files_match_reference(tmp_path, name):
    banner, rows, cols, ent = CASES[name]
    path = write(tmp_path / f"{name}.mtx", banner, rows, cols, ent)
    want = oracle.ref_read_mtx(path)
    got = oracle.read_mtx(path)
    for g, w in zip(got[2:], want):
        assert g.tobytes() == w.tobytes()  # the restatement is the reference
    _, lib = load(path)
    for g, w in zip(lib[1:], want):
        assert g.tobytes() == w.tobytes()


@pytest.mark.parametrize("banner,lower", [("real general", False), ("real symmetric", True),
                                          ("pattern general", False)])
def test_large_file_parallel_parse(tmp_path, banner, lower):
    # ~200k entries (several 64 KB chunks), many repeated coordinates:
    # every thread count gives the same bytes, equal to the reference's.
    rng = np.random.default_rng(11)
    ent = random_entries(rng, 3000, 3000, 200_000, pattern=banner.startswith("pattern"),
                         lower=lower)
    path = write(tmp_path / "big.mtx", banner, 3000, 3000, ent)
    results = [load(path, threads=t)[1] for t in (1, 3, 16)]
    for r in results[1:]:
        for a, b in zip(r[1:], results[0][1:]):
            assert a.tobytes() == b.tobytes()
    _, rp, ci, v = results[0]
    assert np.all(np.diff(rp) >= 0)
    for i in range(0, 3000, 97):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)  # sorted, distinct
    if oracle.ref_mtx_available():
        want = oracle.ref_read_mtx(path)
        for a, b in zip((rp, ci, v), want):
            assert a.tobytes() == b.tobytes()
    # scipy pins the structure and the values to FP64 rounding of the sums.
    import scipy.io
    import scipy.sparse
    S = scipy.sparse.csr_matrix(scipy.io.mmread(path))
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(S.indptr, rp) and np.array_equal(S.indices, ci)
    np.testing.assert_allclose(v, S.data, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 2.0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 2.0\n",  # truncated
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",  # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1.0\n",  # 0 is not 1-based
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 x\n",
    "not a matrix market file\n",
    "%%MatrixMarket matrix coordinate real general\n2 2\n",
])
def test_malformed_files_are_rejected(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    m = F.CsrMatrix()
    assert L.rvsp_csr_read_mtx(str(p).encode(), 0, C.byref(m)) == F.ERROR_INVALID_ARGUMENT
    assert m.row_ptr is None and m.owns_data == 0


def test_argument_checks(tmp_path):
    m = F.CsrMatrix()
    assert L.rvsp_csr_read_mtx(None, 0, C.byref(m)) == F.ERROR_NULL_POINTER
    assert L.rvsp_csr_read_mtx(b"x", 0, None) == F.ERROR_NULL_POINTER
    missing = str(tmp_path / "missing.mtx").encode()
    assert L.rvsp_csr_read_mtx(missing, 0, C.byref(m)) == F.ERROR_INVALID_ARGUMENT


def test_canonicalize_in_place():
    # Unsorted rows with repeats: sorted, folded in order, nnz shrinks.
    rng = np.random.default_rng(5)
    rows, cols = 200, 40
    counts = rng.integers(0, 12, rows)
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    ci = rng.integers(0, cols, int(rp[-1])).astype(np.int32)
    v = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    want_rp, want_ci, want_v = [0], [], []
    for i in range(rows):
        pairs = sorted(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()),
                       key=lambda t: t[0])
        j = 0
        while j < len(pairs):
            c, acc = pairs[j][0], np.float32(pairs[j][1])
            j += 1
            while j < len(pairs) and pairs[j][0] == c:
                acc = np.float32(acc + np.float32(pairs[j][1]))
                j += 1
            want_ci.append(c)
            want_v.append(acc)
        want_rp.append(len(want_ci))
    a_rp, a_ci, a_v = rp.copy(), ci.copy(), v.copy()
    m = F.CsrMatrix(rows, cols, int(rp[-1]), a_rp.ctypes.data, a_ci.ctypes.data, a_v.ctypes.data,
                    F.DTYPE_FP32, F.FORMAT_CSR, 0)
    assert L.rvsp_csr_canonicalize(C.byref(m)) == F.SUCCESS
    n = m.nnz
    assert n == len(want_ci) < int(rp[-1])
    assert a_rp.tolist() == want_rp
    assert a_ci[:n].tolist() == want_ci
    assert a_v[:n].tobytes() == np.array(want_v, np.float32).tobytes()
    # Out-of-range columns are refused with the matrix untouched.
    bad_ci = ci.copy()
    bad_ci[3] = cols
    b_rp = rp.copy()
    mb = F.CsrMatrix(rows, cols, int(rp[-1]), b_rp.ctypes.data, bad_ci.ctypes.data,
                     v.copy().ctypes.data, F.DTYPE_FP32, F.FORMAT_CSR, 0)
    assert L.rvsp_csr_canonicalize(C.byref(mb)) == F.ERROR_INVALID_ARGUMENT
    assert b_rp.tobytes() == rp.tobytes()


def test_restatement_planted_mistake_is_caught(tmp_path):
    # Summing repeated entries in reverse order changes the FP32 bits of a
    # row whose repeats do not associate: the restatement's order matters.
    ent = [(1, 1, "1"), (1, 1, "1e8"), (1, 1, "-1e8")]
    path = write(tmp_path / "order.mtx", "real general", 1, 1, ent)
    _, _, _, _, v = oracle.read_mtx(path)
    fwd = np.float32(np.float32(np.float32(1) + np.float32(1e8)) + np.float32(-1e8))
    rev = np.float32(np.float32(np.float32(-1e8) + np.float32(1e8)) + np.float32(1))
    assert fwd != rev and v[0] == fwd
    st, got = load(path)
    assert got[3][0] == fwd
    if oracle.ref_mtx_available():
        assert oracle.ref_read_mtx(path)[2][0] == fwd


@pytest.mark.gpu
def test_loaded_matrix_multiplies_on_the_gpu(gpu, tmp_path):
    # The reference's bench path: Matrix Market file -> CSR -> C = A * A,
    # here through the library's loader and the HIP product, bit for bit
    # against the oracle.
    import rvsp
    from rvsp.api import Csr
    rng = np.random.default_rng(21)
    ent = random_entries(rng, 2500, 2500, 40_000, lower=True)
    path = write(tmp_path / "sym.mtx", "real symmetric", 2500, 2500, ent)
    st, got = load(path)
    assert st == F.SUCCESS
    (rows, cols, _, _), rp, ci, v = got
    A = Csr(rows, cols, rp, ci, v)
    C_ = rvsp.spgemm(A, A)
    want = oracle.spgemm(A, A)
    for g, w in zip((C_.row_ptr, C_.col_idx, C_.values), want):
        assert g.tobytes() == w.tobytes()
