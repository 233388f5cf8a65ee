"""Values of the HIP library against the FP64 oracle on every golden case
(all five workload shapes, at the sizes the digests were recorded on): the
condition-aware error max |c - c64| / sum_k |a_ik b_kj| must stay within
1e-5. Where no cancellation is possible (non-negative operands: the
Laplacian and R-MAT workloads) that is also the plain relative error. With
signed operands an entry that nearly cancels can carry a larger plain
relative error (FP32 summation in the reference's order: the oracle's FP32
result, the reference's bits, shows the same figure, which is asserted).
The FP32 bits themselves are pinned to the reference by the digest tests;
this bounds how far they are from the exact product."""
import numpy as np
import pytest

import cases
import rvsp
from oracle import oracle
from rvsp import _ffi as F

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("name", cases.GOLDEN_CASES)
def test_fp64_oracle_error_within_tolerance(gpu, name):
    A, B = cases.build(name)
    got, _ = rvsp.spgemm_device(A, B, algo=F.ALGO_HIP_LEARNED)
    e = oracle.f64_error(A, B, got.row_ptr, got.col_idx, got.values)
    print(f"{name}: {e}")
    assert e["entries"] == got.nnz
    assert e["max_cond_err"] <= TOL, e
    if np.all(A.values >= 0) and np.all(B.values >= 0):
        assert e["max_rel_err"] <= TOL, e
    rp, ci, v = oracle.spgemm(A, B)
    assert oracle.f64_error(A, B, rp, ci, v) == e
