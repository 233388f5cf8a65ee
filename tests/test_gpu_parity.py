"""GPU parity: the HIP library against the pinned oracle and the golden
digests recorded from the reference library. Structure and values must be
bit-identical (the GPU accumulates each C(i,j) in the reference's order
with separately rounded multiply and add)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import cases
import rvsp
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import Csr

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))["fixtures"]
DIGESTS = json.load(open(os.path.join(HERE, "golden", "digests.json")))["cases"]


def fx_csr(m):
    return Csr(m["rows"], m["cols"], m["row_ptr"], m["col_idx"], m["values"])


def same(got, want):
    rp, ci, val = want
    assert got.row_ptr.tobytes() == rp.tobytes(), "row_ptr differs"
    assert got.col_idx.tobytes() == ci.tobytes(), "col_idx differs"
    if got.values.tobytes() != val.tobytes():
        bad = np.nonzero(got.values.view(np.uint32) != val.view(np.uint32))[0]
        raise AssertionError(f"{bad.size} values differ, first at {bad[:5]}: "
                             f"{got.values[bad[:5]]} vs {val[bad[:5]]}")


@pytest.mark.parametrize("fx", FIXTURES, ids=[f["name"] for f in FIXTURES])
def test_hand_fixtures_all_entry_points(gpu, fx):
    A, B, want = fx_csr(fx["a"]), fx_csr(fx["b"]), fx_csr(fx["c"])
    exp = (want.row_ptr, want.col_idx, want.values)
    for algo in (F.ALGO_DEFAULT, F.ALGO_HIP_FIXED, F.ALGO_HIP_TUNED, F.ALGO_HIP_LEARNED):
        same(rvsp.spgemm(A, B, algo=algo), exp)
    same(rvsp.spgemm_device(A, B)[0], exp)
    for backend in (F.BACKEND_SCALAR, F.BACKEND_SCALAR_UNROLL4, F.BACKEND_HIP):
        same(rvsp.spgemm_oneshot(A, B, backend=backend), exp)


def test_cancelled_entry_is_positive_zero(gpu):
    fx = next(f for f in FIXTURES if f["name"] == "cancel")
    C_ = rvsp.spgemm(fx_csr(fx["a"]), fx_csr(fx["b"]))
    assert C_.values.view(np.uint32).tolist() == [0]


@pytest.mark.parametrize("name", cases.GOLDEN_CASES)
def test_golden_digest_host_mode(gpu, name):
    A, B = cases.build(name)
    got = rvsp.spgemm(A, B)
    assert oracle.digest(got.row_ptr, got.col_idx, got.values) == DIGESTS[name]["digest"]


@pytest.mark.parametrize("name", cases.GOLDEN_CASES)
def test_golden_digest_device_mode(gpu, name):
    A, B = cases.build(name)
    got, plan = rvsp.spgemm_device(A, B)
    assert oracle.digest(got.row_ptr, got.col_idx, got.values) == DIGESTS[name]["digest"]


def _random_pair(seed):
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    m, k, n = (int(x) for x in rng.integers(1, 5000, size=3))
    lam_a, lam_b = rng.uniform(0.2, 30, size=2)
    ca = np.minimum(rng.poisson(lam_a, size=m), k)
    cb = np.minimum(rng.poisson(lam_b, size=k), n)
    if seed % 3 == 0:
        ca[rng.integers(0, m, size=3)] = min(k, 400)
    rp, ci = gen.random_rows(m, k, ca, rng)
    A = Csr(m, k, rp, ci, rng.standard_normal(ci.size))
    rp, ci = gen.random_rows(k, n, cb, rng)
    B = Csr(k, n, rp, ci, rng.standard_normal(ci.size))
    return A, B


@pytest.mark.parametrize("seed", range(12))
def test_random_shapes_bit_exact(gpu, seed):
    A, B = _random_pair(seed)
    want = oracle.spgemm(A, B)
    same(rvsp.spgemm(A, B), want)
    same(rvsp.spgemm_device(A, B)[0], want)


# Forced plans: every accumulator kind and lane count a bin can take must give
# the same bits; impossible choices fall back to global dense.
PLAN_VARIANTS = {
    "all_global": [(ph, b, F.KIND_GLOBAL, 0, 0) for ph in (0, 1, 2) for b in range(2, 24)],
    "numeric_block": [(2, b, F.KIND_BLOCK, 256, 4096) for b in range(2, 12)],
    # Small tiles: rows span many tiles (more than the 16 of the launch grid),
    # so long B rows are narrowed and tiles loop.
    "numeric_block_t256": [(2, b, F.KIND_BLOCK, 256, 256) for b in range(2, 13)],
    "numeric_block_t1024": [(2, b, F.KIND_BLOCK, 256, 1024) for b in range(2, 13)],
    # Wave tiles (KIND_BLOCK variant 1): 256-position tiles split rows into
    # many runs and tiles (narrowing, cursors, hash-indexed sparse tiles);
    # 2048 leaves most rows whole.
    "numeric_wave_t256": [(2, b, F.KIND_BLOCK, 256, 256, 1) for b in range(2, 13)],
    "numeric_wave_t1024": [(2, b, F.KIND_BLOCK, 256, 1024, 1) for b in range(2, 13)],
    "numeric_wave_t2048": [(2, b, F.KIND_BLOCK, 256, 2048, 1) for b in range(2, 13)],
    "numeric_search_g1": [(2, b, F.KIND_SEARCH, 1, 0) for b in range(2, 6)],
    "numeric_search_g64": [(2, b, F.KIND_SEARCH, 64, 0) for b in range(2, 10)],
    "hash_g2": [(ph, b, F.KIND_HASH, 2, 0) for ph in (0, 1) for b in range(2, 7)],
    "too_small_tables": [(ph, b, 0, 0, 4) for ph in (0, 1, 2) for b in range(2, 10)],
    "direct_g1": [(ph, 1, F.KIND_DIRECT, 1, 0) for ph in (0, 1, 2)],
    "numeric_map_g1": [(2, b, F.KIND_MAP, 1, 0) for b in range(2, 5)],
    "numeric_map_g4": [(2, b, F.KIND_MAP, 4, 0) for b in range(2, 10)],
    "numeric_map_g64": [(2, b, F.KIND_MAP, 64, 0) for b in range(2, 10)],
    "numeric_map_staged_a": [(2, b, F.KIND_MAP, 2, 0, 1) for b in range(2, 10)],
    "numeric_map_chunked": [(2, b, F.KIND_MAP, 1, 0, 2) for b in range(2, 10)],
}


@pytest.mark.parametrize("name", ["cfg1_uniform16_n4096", "cfg2_laplacian_g1024_it0",
                                  "cfg3_rmat_s12", "cfg5_banded_small"])
def test_position_map_repeated_numeric(gpu, name):
    # The map is built on the first compute and drives every later one:
    # redraw the values, recompute numeric-only, compare with the oracle.
    A, B = cases.build(name)
    choices = [(2, b, F.KIND_MAP, 2, 0, v) for b in range(2, 10) for v in (0, 1)][1::2] \
        if name.startswith("cfg2") else [(2, b, F.KIND_MAP, 2, 0) for b in range(2, 10)]
    dA = rvsp.DeviceCsr.upload(A)
    dB = dA if B is A else rvsp.DeviceCsr.upload(B)
    d = rvsp.Descriptor(memory=F.MEMORY_DEVICE, choices=choices)
    nnz = d.estimate(dA, dB)
    assert any(p["kind"] == "lds_map" and p["rows"] > 0 for p in d.plan())
    dC = rvsp.DeviceCsr(A.rows, B.cols, nnz)
    Cm = dC.struct(nnz)
    rng = np.random.Generator(np.random.PCG64(77))
    square = B is A
    for it in range(3):
        if it:
            A = A.with_values(rng.uniform(-1, 1, A.nnz).astype(np.float32))
            dA.set_values(A.values)
            if square:
                B = A
        F.check(d.compute(dA, dB, Cm))
        same(dC.download(), oracle.spgemm(A, B))
    d.close()
    dC.free()
    dA.free()
    if dB is not dA:
        dB.free()


@pytest.mark.parametrize("name", ["cfg3_rmat_s12", "long_rows"])
def test_tiled_bins_move_to_the_map_on_reuse(gpu, name):
    # Tiled numeric bins of rows up to 2048 long move to the position map on
    # the first numeric-only compute; every compute keeps the oracle's bits.
    A, B = cases.build(name)
    choices = [(2, b, F.KIND_BLOCK, 256, 1024) for b in range(2, 10)]
    dA = rvsp.DeviceCsr.upload(A)
    dB = dA if B is A else rvsp.DeviceCsr.upload(B)
    d = rvsp.Descriptor(memory=F.MEMORY_DEVICE, choices=choices)
    nnz = d.estimate(dA, dB)
    before = [p for p in d.plan() if p["phase"] == 2 and p["rows"] > 0 and 2 <= p["bin"] < 10]
    assert before and all(p["kind"] == "lds_block" for p in before)
    dC = rvsp.DeviceCsr(A.rows, B.cols, nnz)
    Cm = dC.struct(nnz)
    rng = np.random.Generator(np.random.PCG64(5))
    square = B is A
    for it in range(3):
        if it:
            A = A.with_values(rng.uniform(-1, 1, A.nnz).astype(np.float32))
            dA.set_values(A.values)
            if square:
                B = A
        F.check(d.compute(dA, dB, Cm))
        same(dC.download(), oracle.spgemm(A, B))
    after = [p for p in d.plan() if p["phase"] == 2 and p["rows"] > 0 and 2 <= p["bin"] < 10]
    assert all(p["kind"] == "lds_map" for p in after)
    d.close()
    dC.free()
    dA.free()
    if dB is not dA:
        dB.free()


@pytest.mark.parametrize("variant", sorted(PLAN_VARIANTS))
@pytest.mark.parametrize("name", ["cfg4_pair_0", "long_rows", "rect_empty_rows",
                                  "cfg2_laplacian_g64_stencil", "cfg3_rmat_s12"])
def test_forced_plans_identical_bits(gpu, name, variant):
    A, B = cases.build(name)
    got, plan = rvsp.spgemm_device(A, B, choices=PLAN_VARIANTS[variant])
    assert oracle.digest(got.row_ptr, got.col_idx, got.values) == DIGESTS[name]["digest"]


@pytest.mark.parametrize("cap", [512, 1024])
def test_wave_tiles_on_power_law_digest(gpu, cap):
    # R-MAT 2^16 (hub rows of thousands of entries, spans of the whole
    # column range): every long-row numeric bin on wave tiles.
    name = "cfg3_rmat_s16_seed3"
    A, B = cases.build(name)
    choices = [(2, b, F.KIND_BLOCK, 256, cap, 1) for b in range(10, 13)]
    got, plan = rvsp.spgemm_device(A, B, choices=choices)
    assert any(p["phase"] == 2 and p["kind"] == "lds_block" and p["rows"] > 0 for p in plan)
    assert oracle.digest(got.row_ptr, got.col_idx, got.values) == DIGESTS[name]["digest"]


def test_plan_uses_several_kernel_kinds(gpu):
    A, B = cases.build("long_rows")
    _, plan = rvsp.spgemm_device(A, B)
    used = {(p["phase"], p["kind"]) for p in plan if p["rows"] > 0}
    kinds = {k for _, k in used}
    assert {"direct", "lds_hash", "lds_search"} <= kinds
    assert any(k in kinds for k in ("lds_block", "lds_bitmap", "global_dense"))


def test_op_counts_match_oracle(gpu):
    A, B = cases.build("cfg3_rmat_s12")
    d = rvsp.Descriptor()
    d.estimate(A, B)
    _, ops = oracle.symbolic(A, B)
    assert d.op_counts(A.rows).tolist() == ops.tolist()
    d.close()


def test_descriptor_reuse_numeric_only(gpu):
    # cfg2 in miniature: analyse once, then numeric-only repeats with values
    # redrawn each time; every repeat must equal the oracle on those values.
    A = gen.laplacian_2d(96)
    dA = rvsp.DeviceCsr.upload(A)
    d = rvsp.Descriptor(memory=F.MEMORY_DEVICE)
    nnz = d.estimate(dA, dA)
    dC = rvsp.DeviceCsr(A.rows, A.cols, nnz)
    Cm = dC.struct(nnz)
    for it in range(4):
        vals = gen.laplacian_values(A, it)
        dA.set_values(vals)
        F.check(d.compute(dA, dA, Cm))
        A_it = A.with_values(vals)
        same(dC.download(), oracle.spgemm(A_it, A_it))
    d.close()
    dC.free()
    dA.free()


def test_reuse_host_mode_new_buffers(gpu):
    # A new C.col_idx buffer gets the structure again; the same buffer keeps it.
    A, B = cases.build("cfg4_pair_2")
    d = rvsp.Descriptor()
    nnz = d.estimate(A, B)
    want = oracle.spgemm(A, B)
    for _ in range(2):
        rp = np.zeros(A.rows + 1, np.int32)
        ci = np.full(nnz, -5, np.int32)
        v = np.zeros(nnz, np.float32)
        Cm = F.CsrMatrix(A.rows, B.cols, nnz, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                         F.DTYPE_FP32, F.FORMAT_CSR, 0)
        F.check(d.compute(A, B, Cm))
        same(Csr(A.rows, B.cols, rp, ci, v), want)
    ci[:] = -7
    F.check(d.compute(A, B, Cm))
    assert np.all(ci == -7)  # same buffer: structure not rewritten
    d.invalidate()
    F.check(d.compute(A, B, Cm))
    assert ci.tobytes() == want[1].tobytes()
    d.close()


def test_int32_overflow_reports_allocation_failed(gpu):
    # 65536 rows each selecting one B row of 40000 entries: nnz(C) = 2.6e9.
    m, width = 65536, 40000
    A = Csr(m, 1, np.arange(m + 1), np.zeros(m), np.ones(m))
    B = Csr(1, width, [0, width], np.arange(width), np.ones(width))
    d = rvsp.Descriptor()
    with pytest.raises(rvsp.RvspError) as e:
        d.estimate(A, B)
    assert e.value.status == F.ERROR_ALLOCATION_FAILED
    d.close()


def test_raw_oneshot_kernel(gpu):
    A, B = cases.build("cfg4_pair_3")
    L = F.lib()
    need = C.c_size_t()
    L.rvsp_spgemm_buffer_size(B.cols, C.byref(need))
    ws = np.zeros(need.value, np.uint8)
    rp, ci, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int32()
    ops = np.zeros(A.rows, np.int64)
    st = L.rvsp_spgemm_hip_f32(A.rows, A.cols, B.cols, A.row_ptr.ctypes.data,
                               A.col_idx.ctypes.data, A.values.ctypes.data,
                               B.row_ptr.ctypes.data, B.col_idx.ctypes.data,
                               B.values.ctypes.data, ws.ctypes.data, need.value,
                               C.byref(rp), C.byref(ci), C.byref(val), C.byref(nnz),
                               ops.ctypes.data)
    assert st == F.SUCCESS
    n = nnz.value
    got = Csr(A.rows, B.cols,
              np.ctypeslib.as_array((C.c_int32 * (A.rows + 1)).from_address(rp.value)).copy(),
              np.ctypeslib.as_array((C.c_int32 * n).from_address(ci.value)).copy(),
              np.ctypeslib.as_array((C.c_float * n).from_address(val.value)).copy())
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    for p in (rp, ci, val):
        libc.free(p)
    same(got, oracle.spgemm(A, B))
    assert ops.tolist() == oracle.symbolic(A, B)[1].tolist()
    assert L.rvsp_spgemm_hip_f32(A.rows, A.cols, B.cols, A.row_ptr.ctypes.data,
                                 A.col_idx.ctypes.data, A.values.ctypes.data,
                                 B.row_ptr.ctypes.data, B.col_idx.ctypes.data,
                                 B.values.ctypes.data, ws.ctypes.data, need.value - 1,
                                 C.byref(rp), C.byref(ci), C.byref(val), C.byref(nnz),
                                 None) == F.ERROR_INVALID_ARGUMENT


def test_full_size_banded_with_dense_rows(gpu):
    # The banded workload at full size: A 500k x 200k (every 100th row 2000
    # uniform entries over 200k columns), B 200k x 500k: C's dense rows span
    # most of B's 500k columns (long-row kernels). FIXED and LEARNED plans give
    # the same bytes; oracle rows cover band rows and dense rows at both ends.
    A, B = gen.banded_mixed(seed=0)
    got_f, _ = rvsp.spgemm_device(A, B, algo=F.ALGO_HIP_FIXED)
    got_l, plan = rvsp.spgemm_device(A, B, algo=F.ALGO_HIP_LEARNED)
    for k in ("row_ptr", "col_idx", "values"):
        assert getattr(got_f, k).tobytes() == getattr(got_l, k).tobytes()
    assert any(p["phase"] == 2 and p["kind"] == "lds_block" and p["rows"] > 0 for p in plan)
    for r0, r1 in ((0, 3000), (A.rows - 2100, A.rows)):
        rp, ci, v = oracle.spgemm(A, B, r0, r1)
        s, e = int(got_f.row_ptr[r0]), int(got_f.row_ptr[r1])
        assert np.array_equal(got_f.row_ptr[r0:r1 + 1] - s, rp)
        assert got_f.col_idx[s:e].tobytes() == ci.tobytes()
        assert got_f.values[s:e].tobytes() == v.tobytes()
