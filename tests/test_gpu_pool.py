"""The library's device buffer cache: released descriptor state is reused by
later descriptors (the test suite runs with RVSP_POOL_POISON=1, so every
reused buffer arrives filled with 0xFF bytes) and results stay bit-exact."""
import numpy as np
import pytest

import rvsp
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen

pytestmark = pytest.mark.gpu
L = F.lib()


def product(A, B, algo):
    got, _ = rvsp.spgemm_device(A, B, algo=algo)
    return got


def same(got, want):
    rp, ci, v = want
    return (got.row_ptr.tobytes() == rp.tobytes() and got.col_idx.tobytes() == ci.tobytes()
            and got.values.tobytes() == v.tobytes())


def test_reused_buffers_give_the_same_bits(gpu):
    assert L.rvsp_hip_pool_release() == F.SUCCESS
    assert L.rvsp_hip_pool_cached_bytes() == 0
    A = gen.uniform_k(3000, 14, seed=11)
    R = gen.rmat(13, seed=3)
    want_a = oracle.spgemm(A, A)
    want_r = oracle.spgemm(R, R)
    for algo in (F.ALGO_HIP_FIXED, F.ALGO_HIP_LEARNED):
        assert same(product(A, A, algo), want_a)
        cached = L.rvsp_hip_pool_cached_bytes()
        assert cached > 0, "a destroyed descriptor's buffers stay cached"
        # The next products take (poisoned) cached buffers of other shapes.
        assert same(product(R, R, algo), want_r)
        assert same(product(A, A, algo), want_a)
    assert L.rvsp_hip_pool_release() == F.SUCCESS
    assert L.rvsp_hip_pool_cached_bytes() == 0
    assert same(product(R, R, F.ALGO_HIP_FIXED), want_r)
