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


def test_autotuned_bin_thresholds_are_recorded_and_applied(gpu, tmp_path):
    # The tuner also picks each phase's bin thresholds (rule variants) and
    # records them under bin 0 of the table; a HIP_TUNED estimation re-bins
    # with them. Forced here: a table whose numeric rule stops the searching
    # bins at 256-column rows must leave the bins beyond 256 empty and still
    # give the oracle's bits.
    A, B = cases.build("long_rows")
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
    d.close()
    path = tmp_path / "table.txt"
    assert L.rvsp_tuning_table_save(str(path).encode()) > 0
    rows = [l.split() for l in path.read_text().splitlines() if not l.startswith("#")]
    rules = {int(r[0]): int(r[5]) for r in rows if int(r[1]) == 0}
    assert sorted(rules) == [0, 1, 2]  # one rule per phase
    assert rules[0] in (2, 3, 4, 5, 6) and rules[1] in (2, 3, 4)
    assert rules[2] in (256, 512, 1024, 2048)
    # Force the narrowest numeric rule and check the plan and the bits.
    key = max(0, min(10, int(np.floor(np.log2(max(1.0, B.nnz / B.rows))))))
    text = "".join(" ".join(r) + "\n" for r in rows if not (int(r[1]) == 0 and int(r[0]) == 2))
    text += f"2 0 {key} 0 0 256 1.0 0\n"
    path.write_text(text)
    L.rvsp_tuning_table_clear()
    assert L.rvsp_tuning_table_load(str(path).encode()) > 0
    got, plan = rvsp.spgemm_device(A, B, algo=F.ALGO_HIP_TUNED)
    for g, w in zip((got.row_ptr, got.col_idx, got.values), want):
        assert g.tobytes() == w.tobytes()
    search = [p for p in plan if p["phase"] == 2 and p["kind"] in ("lds_search", "lds_map")]
    assert search and all(p["rows"] == 0 for p in search if p["cap"] > 256)
    L.rvsp_tuning_table_clear()
