"""The canonical-CSR oracle against the reference's own rvsp_csr_check
(compiled from its sources into oracle/_ref) and against the library's host
check, on seeded structures with planted faults. No device needed."""
import ctypes as C

import numpy as np
import pytest

import cases
from oracle import oracle
from rvsp import _ffi as F

L = F.lib()
CASES = cases.csr_check_cases()


def host_check(rows, cols, rp, ci):
    rpa = np.ascontiguousarray(rp, np.int32)
    cia = None if ci is None else np.ascontiguousarray(ci if len(ci) else [0], np.int32)
    bad = C.c_int32(-7)
    st = L.rvsp_csr_check(rows, cols, rpa.ctypes.data, None if cia is None else cia.ctypes.data,
                          C.byref(bad))
    return st, bad.value


def test_cases_cover_every_verdict():
    verdicts = {oracle.csr_check(r, c, rp, ci)[0] for _, r, c, rp, ci in CASES}
    assert verdicts == {oracle.CSR_OK, oracle.CSR_BAD_ROW_PTR, oracle.CSR_COL_OUT_OF_RANGE,
                        oracle.CSR_UNSORTED, oracle.CSR_DUPLICATE}


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_host_check(case):
    _, rows, cols, rp, ci = case
    assert oracle.csr_check(rows, cols, rp, ci) == host_check(rows, cols, rp, ci)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built")
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference(case):
    _, rows, cols, rp, ci = case
    assert oracle.csr_check(rows, cols, rp, ci) == oracle.ref_csr_check(rows, cols, rp, ci)


@pytest.mark.parametrize("variant", [1, 2])
def test_planted_oracle_mistakes_are_caught(variant):
    # A restatement that let duplicates pass, or kept the last failure
    # instead of the first, must disagree with the library's host check
    # (itself matched to the reference above) on some case.
    wrong = [c[0] for c in CASES
             if oracle.csr_check(c[1], c[2], c[3], c[4], variant=variant) != host_check(*c[1:])]
    assert wrong
