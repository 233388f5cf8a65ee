"""The legacy-dtype oracle (FP64 -> FP64, INT8 -> INT32) against the
reference's own legacy kernels (oracle/_ref) and against exact dense
products; a planted order mistake must be caught. No device needed."""
import numpy as np
import pytest

import cases
from oracle import oracle
from rvsp.api import Csr


def dense(M, dt):
    out = np.zeros((M.rows, M.cols), dtype=dt)
    out[M.row_ids(), M.col_idx] = M.values
    return out


def exact_dense_product(A, B, i8):
    """Integer-valued operands: the exact product (int64, wrapped to int32
    for INT8; FP64 sums of small integers are exact)."""
    if i8:
        prod = dense(A, np.int64) @ dense(B, np.int64)
        return ((prod + 2**31) % 2**32 - 2**31).astype(np.int32)
    return dense(A, np.float64) @ dense(B, np.float64)


def as_dense(rp, ci, val, rows, cols, dt):
    out = np.zeros((rows, cols), dtype=dt)
    out[np.repeat(np.arange(rows), np.diff(rp)), ci] = val
    return out


@pytest.mark.parametrize("i8", [False, True], ids=["f64", "i8"])
@pytest.mark.parametrize("seed", range(6))
def test_legacy_oracle_matches_reference(i8, seed):
    if not oracle.ref_available():
        pytest.skip("reference library not built")
    A, B = cases.legacy_pair(seed, i8)
    got = oracle.spgemm_legacy(A, B)
    st, want = oracle.ref_spgemm_legacy(A, B)
    assert st == 0
    for g, w in zip(got, want):
        assert g.dtype == w.dtype and g.tobytes() == w.tobytes()


@pytest.mark.parametrize("i8", [False, True], ids=["f64", "i8"])
def test_legacy_oracle_exact_on_integer_values(i8):
    rng = np.random.default_rng(3)
    for trial in range(8):
        A, B = cases.legacy_pair(100 + trial, i8, rows=40, inner=30, cols=200)
        if not i8:  # integer values only: every sum is exact in FP64
            B = Csr(B.rows, B.cols, B.row_ptr, B.col_idx,
                    rng.integers(-3, 4, size=B.nnz).astype(np.float64), dtype=np.float64)
        rp, ci, val = oracle.spgemm_legacy(A, B)
        want = exact_dense_product(A, B, i8)
        got = as_dense(rp, ci, val, A.rows, B.cols, want.dtype)
        assert np.array_equal(got, want)
        # Exactly the non-zero sums are stored, in ascending columns.
        assert int(rp[-1]) == int(np.count_nonzero(want))
        assert np.all(val != 0)
        for r in range(A.rows):
            assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)


def test_legacy_cancellation_drops_entries():
    A = Csr(1, 2, [0, 2], [0, 1], [1.0, 1.0], dtype=np.float64)
    B = Csr(2, 3, [0, 2, 4], [0, 2, 0, 1], [2.0, 1.0, -2.0, 5.0], dtype=np.float64)
    rp, ci, val = oracle.spgemm_legacy(A, B)  # column 0 cancels: 2 - 2
    assert rp.tolist() == [0, 2] and ci.tolist() == [1, 2] and val.tolist() == [5.0, 1.0]


def test_planted_order_mistake_is_caught():
    # A's entries taken in descending order give different FP64 bits on
    # fractional values: the reference comparison must see it.
    if not oracle.ref_available():
        pytest.skip("reference library not built")
    differs = False
    for seed in range(6):
        A, B = cases.legacy_pair(seed, False)
        bad = oracle.spgemm_legacy(A, B, variant=1)
        _, want = oracle.ref_spgemm_legacy(A, B)
        differs |= any(g.tobytes() != w.tobytes() for g, w in zip(bad, want))
    assert differs
