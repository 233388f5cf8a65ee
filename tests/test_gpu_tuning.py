"""Autotuner and tuned plans: tuning changes kernel shapes only, never bits."""
import ctypes as C
import os

import numpy as np
import pytest

import cases
import rvsp
from oracle import oracle
from rvsp import _ffi as F

pytestmark = pytest.mark.gpu
L = F.lib()


@pytest.mark.parametrize("name", ["cfg2_laplacian_g64_stencil", "cfg4_pair_0", "long_rows"])
def test_autotune_keeps_bits_and_feeds_tuned_plans(gpu, name, tmp_path):
    A, B = cases.build(name)
    want = oracle.spgemm(A, B)
    L.rvsp_tuning_table_clear()
    d = rvsp.Descriptor(algo=F.ALGO_HIP_TUNED)
    nnz = d.estimate(A, B)
    rp = np.zeros(A.rows + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    v = np.zeros(max(nnz, 1), np.float32)
    Cm = F.CsrMatrix(A.rows, B.cols, nnz, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                     F.DTYPE_FP32, F.FORMAT_CSR, 0)
    a, b = A.struct(), B.struct()
    F.check(L.rvsp_spgemm_autotune(d.handle, C.byref(a), C.byref(b), C.byref(Cm),
                                   C.c_void_p(d.ws.ctypes.data), 2), "autotune")
    for g, w in zip((rp, ci[:nnz], v[:nnz]), want):
        assert g.tobytes() == w.tobytes()
    # A numeric-only compute with the tuned plan still matches.
    F.check(d.compute(A, B, Cm))
    assert v[:nnz].tobytes() == want[2].tobytes()
    d.close()
    n = L.rvsp_tuning_table_size()
    assert n > 0
    path = str(tmp_path / "table.txt").encode()
    assert L.rvsp_tuning_table_save(path) == n
    L.rvsp_tuning_table_clear()
    assert L.rvsp_tuning_table_load(path) == n
    # A fresh HIP_TUNED descriptor picks the table up and gives the same bits.
    got = rvsp.spgemm(A, B, algo=F.ALGO_HIP_TUNED)
    assert got.values.tobytes() == want[2].tobytes()
    L.rvsp_tuning_table_clear()
