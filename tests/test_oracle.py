"""Pins the CPU oracle (oracle/spgemm_oracle.c) before anything is checked
against it: the reference's hand fixtures, golden digests recorded from the
reference library, the reference library itself where it is built, brute
force on tiny inputs, a closed form, and planted mistakes that must fail."""
import json
import os

import numpy as np
import pytest

import cases
from oracle import oracle
from rvsp import gen
from rvsp.api import Csr

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["fixtures"]
DIGESTS = json.load(open(os.path.join(HERE, "golden", "digests.json")))["cases"]


def fx_csr(m):
    return Csr(m["rows"], m["cols"], m["row_ptr"], m["col_idx"], m["values"])


@pytest.mark.parametrize("fx", FIXTURES, ids=[f["name"] for f in FIXTURES])
def test_oracle_hand_fixtures(fx):
    # Reference tests/test_spgemm_reference_f32.c:131-148: structure exact;
    # the expected values are small integers, so they must match exactly.
    A, B, want = fx_csr(fx["a"]), fx_csr(fx["b"]), fx_csr(fx["c"])
    rp, ci, val = oracle.spgemm(A, B)
    assert rp.tolist() == want.row_ptr.tolist()
    assert ci.tolist() == want.col_idx.tolist()
    assert val.tolist() == want.values.tolist()
    if "op_counts" in fx:
        _, ops = oracle.symbolic(A, B)
        assert ops.tolist() == fx["op_counts"]


def test_cancel_fixture_keeps_structural_zero():
    fx = next(f for f in FIXTURES if f["name"] == "cancel")
    rp, ci, val = oracle.spgemm(fx_csr(fx["a"]), fx_csr(fx["b"]))
    assert rp.tolist() == [0, 1] and ci.tolist() == [0]
    assert val[0] == 0.0 and not np.signbit(val[0])


FAST_GOLDEN = [c for c in cases.GOLDEN_CASES if c not in ("cfg3_rmat_s16_seed3",)]


@pytest.mark.parametrize("name", FAST_GOLDEN)
def test_oracle_matches_reference_digest(name):
    A, B = cases.build(name)
    g = DIGESTS[name]
    assert [A.rows, A.cols, A.nnz] == g["a"] and [B.rows, B.cols, B.nnz] == g["b"]
    rp, ci, val = oracle.spgemm(A, B)
    assert int(rp[-1]) == g["nnz_c"]
    assert oracle.digest(rp, ci, val) == g["digest"]
    for s in g["samples"]:
        r = s["row"]
        seg = slice(int(rp[r]), int(rp[r]) + len(s["col_idx"]))
        assert ci[seg].tolist() == s["col_idx"]
        assert val[seg].view(np.uint32).tolist() == s["value_bits"]


@pytest.mark.slow
def test_oracle_matches_reference_digest_rmat16():
    name = "cfg3_rmat_s16_seed3"
    A, B = cases.build(name)
    rp, ci, val = oracle.spgemm(A, B)
    assert oracle.digest(rp, ci, val) == DIGESTS[name]["digest"]


def test_laplacian_nnz_closed_form():
    for g in (3, 8, 64, 1024):
        A = gen.laplacian_2d(g)
        rp, _ = oracle.symbolic(A, A)
        assert int(rp[-1]) == cases.laplacian_squared_nnz(g)
    assert cases.laplacian_squared_nnz(1024) == 13611012


@pytest.mark.parametrize("seed", range(6))
def test_oracle_structure_brute_force(seed):
    rng = np.random.Generator(np.random.PCG64(100 + seed))
    m, k, n = rng.integers(1, 40, size=3)
    ca = np.minimum(rng.poisson(2.0, size=m), k)
    cb = np.minimum(rng.poisson(3.0, size=k), n)
    rp, ci = gen.random_rows(m, k, ca, rng)
    A = Csr(m, k, rp, ci, rng.uniform(-1, 1, ci.size))
    rp, ci = gen.random_rows(k, n, cb, rng)
    B = Csr(k, n, rp, ci, rng.uniform(-1, 1, ci.size))
    orp, oci, _ = oracle.spgemm(A, B)
    brp, bci = oracle.dense_product_structure(A, B)
    assert orp.tolist() == brp.tolist() and oci.tolist() == bci.tolist()
    _, ops = oracle.symbolic(A, B)
    want_ops = [int(sum(B.row_ptr[c + 1] - B.row_ptr[c] for c in A.col_idx[A.row_ptr[i]:A.row_ptr[i + 1]]))
                for i in range(m)]
    assert ops.tolist() == want_ops


@pytest.mark.parametrize("seed", range(4))
def test_oracle_f64_matches_dense_numpy(seed):
    rng = np.random.Generator(np.random.PCG64(200 + seed))
    A = Csr.from_dense(np.where(rng.random((30, 25)) < 0.2, rng.uniform(-1, 1, (30, 25)), 0))
    B = Csr.from_dense(np.where(rng.random((25, 35)) < 0.2, rng.uniform(-1, 1, (25, 35)), 0))
    rp, ci, v32 = oracle.spgemm(A, B)
    v64, mag = oracle.spgemm_f64(A, B, rp, ci)
    dense = A.to_dense().astype(np.float64) @ B.to_dense().astype(np.float64)
    r = np.repeat(np.arange(A.rows), np.diff(rp))
    np.testing.assert_allclose(v64, dense[r, ci], rtol=1e-12, atol=1e-12)
    # FP32 result against FP64, relative to the magnitude of the summands.
    assert np.all(np.abs(v32 - v64) <= 1e-6 * np.maximum(mag, 1e-30))


def test_f64_error_report_against_dense_numpy():
    # The report's two figures recomputed from a dense float64 product and
    # |A| |B|; then a planted error of known size must show up exactly.
    rng = np.random.Generator(np.random.PCG64(77))
    A = Csr.from_dense(np.where(rng.random((40, 30)) < 0.3, rng.uniform(-1, 1, (40, 30)), 0))
    B = Csr.from_dense(np.where(rng.random((30, 50)) < 0.3, rng.uniform(-1, 1, (30, 50)), 0))
    rp, ci, v32 = oracle.spgemm(A, B)
    r = np.repeat(np.arange(A.rows), np.diff(rp))
    Ad, Bd = A.to_dense().astype(np.float64), B.to_dense().astype(np.float64)
    exact = (Ad @ Bd)[r, ci]
    mag = (np.abs(Ad) @ np.abs(Bd))[r, ci]
    e = oracle.f64_error(A, B, rp, ci, v32)
    d = np.abs(v32.astype(np.float64) - exact)
    assert e["entries"] == ci.size
    np.testing.assert_allclose(e["max_cond_err"], (d / mag).max(), rtol=1e-6, atol=1e-15)
    keep = np.abs(exact) >= 1e-6 * mag
    assert e["cancelled_entries"] == int((~keep).sum())
    np.testing.assert_allclose(e["max_rel_err"], (d[keep] / np.abs(exact[keep])).max(),
                               rtol=1e-6, atol=1e-15)
    assert e["max_cond_err"] < 1e-6
    bad = v32.copy()
    k = int(np.argmax(mag))
    bad[k] = np.float32(exact[k] + 1e-3 * mag[k])
    e2 = oracle.f64_error(A, B, rp, ci, bad)
    np.testing.assert_allclose(e2["max_cond_err"], 1e-3, rtol=1e-3)
    m = oracle.merge_f64_errors([e, e2, None])
    assert m["entries"] == 2 * ci.size and m["max_cond_err"] == e2["max_cond_err"]


@pytest.mark.parametrize("variant", [1, 2])
def test_planted_mistakes_are_caught(variant):
    # A fused multiply-add (1) or a reversed A-row order (2) must change the
    # bits on the signed cfg1 product, so the digests pin order and rounding.
    A, B = cases.build("cfg1_uniform16_n4096")
    rp, ci, val = oracle.spgemm(A, B, variant=variant)
    assert oracle.digest(rp, ci, val) != DIGESTS["cfg1_uniform16_n4096"]["digest"]
    # ...while the structure and approximate values are still right.
    _, _, good = oracle.spgemm(A, B)
    np.testing.assert_allclose(val, good, rtol=1e-3, atol=1e-5)


ref = pytest.mark.skipif(not oracle.ref_available(),
                         reason="reference library not built here (no reference tree)")


@ref
@pytest.mark.parametrize("seed", range(8))
def test_oracle_bit_identical_to_reference_library(seed):
    rng = np.random.Generator(np.random.PCG64(300 + seed))
    m, k, n = (int(x) for x in rng.integers(1, 3000, size=3))
    ca = np.minimum(rng.poisson(rng.uniform(0.5, 20), size=m), k)
    cb = np.minimum(rng.poisson(rng.uniform(0.5, 20), size=k), n)
    rp, ci = gen.random_rows(m, k, ca, rng)
    A = Csr(m, k, rp, ci, rng.standard_normal(ci.size))
    rp, ci = gen.random_rows(k, n, cb, rng)
    B = Csr(k, n, rp, ci, rng.standard_normal(ci.size))
    got = oracle.spgemm(A, B)
    want = oracle.ref_spgemm(A, B)
    for g, w in zip(got, want):
        assert g.tobytes() == w.tobytes()


@ref
def test_reference_library_reproduces_committed_digests():
    for name in ("cfg1_uniform16_n4096", "cfg4_pair_1", "rect_empty_rows"):
        A, B = cases.build(name)
        assert oracle.digest(*oracle.ref_spgemm(A, B)) == DIGESTS[name]["digest"]


def test_oracle_row_block_equals_full():
    # Row blocks (used for outputs too large for one int32 CSR) concatenate
    # to the full product.
    A, B = cases.build("cfg4_pair_1")
    rp, ci, val = oracle.spgemm(A, B)
    cut = A.rows // 3
    rp1, ci1, v1 = oracle.spgemm(A, B, 0, cut)
    rp2, ci2, v2 = oracle.spgemm(A, B, cut, A.rows)
    assert np.concatenate([rp1, rp2[1:] + rp1[-1]]).tolist() == rp.tolist()
    assert np.concatenate([ci1, ci2]).tobytes() == ci.tobytes()
    assert np.concatenate([v1, v2]).tobytes() == val.tobytes()


@ref
def test_threaded_reference_timing_runs_every_block():
    # The bench's all-cores CPU figure: the reference's descriptors on
    # product-balanced row blocks, one thread each, all finishing.
    from rvsp import _ffi as F
    A = gen.laplacian_2d(64)
    A = A.with_values(gen.laplacian_values(A, 0))
    rate, used = oracle.time_reference_numeric_threads(A, A, 3, 0.05, F.CsrMatrix)
    assert used == 3 and rate > 0
