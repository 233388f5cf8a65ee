"""Batched products (rvsp_spgemm_batch_*): every pair bit-identical to the
oracle, across groups, output layouts, shared B operands, repeated
numeric-only computes replayed from a HIP graph, and the argument checks."""
import ctypes as C

import numpy as np
import pytest

import rvsp
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import Csr
from rvsp.batch import Batch, BatchOutput, spgemm_batch, upload_pairs

pytestmark = pytest.mark.gpu
L = F.lib()


def small_pairs(count=24, seed=0, n_min=16, n_max=2048):
    return gen.batch_pairs(count, n_min=n_min, n_max=n_max, seed=seed)


def assert_pairs_match(pairs, got):
    for p, ((A, B), g) in enumerate(zip(pairs, got)):
        want = oracle.spgemm(A, B)
        for name, x, w in zip(("row_ptr", "col_idx", "values"),
                              (g.row_ptr, g.col_idx, g.values), want):
            assert x.tobytes() == w.tobytes(), f"pair {p} {name}"


@pytest.mark.parametrize("contiguous", [True, False])
def test_batch_matches_oracle_per_pair(gpu, contiguous):
    pairs = small_pairs()
    got, groups = spgemm_batch(pairs, contiguous=contiguous)
    assert groups == 1
    assert_pairs_match(pairs, got)


@pytest.mark.parametrize("contiguous", [True, False])
def test_batch_groups_under_a_product_limit(gpu, contiguous):
    pairs = small_pairs(count=20, seed=1)
    ops = [int(np.diff(B.row_ptr)[A.col_idx].sum()) for A, B in pairs]
    limit = max(max(ops), sum(ops) // 5)
    got, groups = spgemm_batch(pairs, contiguous=contiguous, group_limit=limit)
    assert groups >= 4
    assert_pairs_match(pairs, got)


def test_batch_of_row_blocks_shares_b(gpu):
    """C = A*A as row blocks of A, all against the same B arrays: B is
    stacked once and every block equals the matching rows of the product."""
    A = gen.rmat(12, seed=4)
    cuts = [0, 700, 701, 2500, A.rows]
    blocks = []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        rp = A.row_ptr[r0:r1 + 1] - A.row_ptr[r0]
        s, e = A.row_ptr[r0], A.row_ptr[r1]
        blocks.append(Csr(r1 - r0, A.cols, rp, A.col_idx[s:e], A.values[s:e]))
    pairs = [(blk, A) for blk in blocks]
    got, _ = spgemm_batch(pairs)
    full = oracle.spgemm(A, A)
    rp, ci, v = full
    for (r0, r1), g in zip(zip(cuts[:-1], cuts[1:]), got):
        s, e = rp[r0], rp[r1]
        assert np.array_equal(g.row_ptr, rp[r0:r1 + 1] - s)
        assert g.col_idx.tobytes() == ci[s:e].tobytes()
        assert g.values.tobytes() == v[s:e].tobytes()


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("repeats", [1, 4])
def test_row_blocks_in_groups_share_the_split_table(gpu, fuse, repeats):
    """Row blocks of C = A*A in several groups, every long row through the
    tiled kernel with small tiles (many tiles per row, narrowed B rows): the
    groups share one split table of the stacked B and every block still
    equals the matching rows of the product, whether the groups' long-row
    bins run as one fused launch (several bins and tile sizes across the
    groups, graph replays included) or one launch per bin and group."""
    A = gen.rmat(12, seed=5)
    cuts = [0, 300, 1100, 2600, A.rows]
    blocks = []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        rp = A.row_ptr[r0:r1 + 1] - A.row_ptr[r0]
        s, e = A.row_ptr[r0], A.row_ptr[r1]
        blocks.append(Csr(r1 - r0, A.cols, rp, A.col_idx[s:e], A.values[s:e]))
    pairs = [(blk, A) for blk in blocks]
    ops = [int(np.diff(A.row_ptr)[blk.col_idx].sum()) for blk in blocks]
    choices = [(2, b, F.KIND_BLOCK, 256, 256 if b % 2 else 512) for b in range(2, 13)]
    got, groups = spgemm_batch(pairs, group_limit=max(ops), choices=choices, fuse=fuse,
                               repeats=repeats)
    assert groups >= 3
    rp, ci, v = oracle.spgemm(A, A)
    for (r0, r1), g in zip(zip(cuts[:-1], cuts[1:]), got):
        s, e = rp[r0], rp[r1]
        assert np.array_equal(g.row_ptr, rp[r0:r1 + 1] - s)
        assert g.col_idx.tobytes() == ci[s:e].tobytes()
        assert g.values.tobytes() == v[s:e].tobytes()


def test_concurrent_groups_with_wide_fill_bitmaps(gpu):
    """Row blocks of an R-MAT 2^20 product in two groups (two streams): rows
    whose fill a single descriptor plans as an LDS bitmap over 2^20 columns
    (one workgroup per CU) are filled from global scratch in the batch; every
    block still equals the oracle's rows."""
    A = gen.rmat(20, seed=0)
    cuts = [(400000, 401500), (700000, 701500)]
    blocks = []
    for r0, r1 in cuts:
        s, e = A.row_ptr[r0], A.row_ptr[r1]
        blocks.append(Csr(r1 - r0, A.cols, A.row_ptr[r0:r1 + 1] - s, A.col_idx[s:e],
                          A.values[s:e]))
    # A single descriptor plans a wide LDS-bitmap fill bin for these rows.
    d = rvsp.Descriptor(algo=F.ALGO_HIP_FIXED)
    d.estimate(blocks[0], A)
    wide = [b for b in d.plan() if b["phase"] == 1 and b["kind"] == "lds_bitmap" and
            b["rows"] > 0 and b["lds_bytes"] > 80 * 1024]
    d.close()
    assert wide, "precondition: a wide LDS-bitmap fill bin"
    ops = [int(np.diff(A.row_ptr)[blk.col_idx].sum()) for blk in blocks]
    got, groups = spgemm_batch([(blk, A) for blk in blocks], group_limit=max(ops))
    assert groups == 2
    for (r0, r1), g in zip(cuts, got):
        rp, ci, v = oracle.spgemm(A, A, r0, r1)
        assert np.array_equal(g.row_ptr, rp)
        assert g.col_idx.tobytes() == ci.tobytes()
        assert g.values.tobytes() == v.tobytes()


@pytest.mark.parametrize("graph", [True, False])
def test_repeated_numeric_computes_with_new_values(gpu, graph):
    pairs = small_pairs(count=12, seed=2)
    As, Bs, owned = upload_pairs(pairs)
    bt = Batch(graph=graph)
    out = None
    try:
        nnz = bt.estimate(As, Bs)
        out = BatchOutput([a.rows for a, _ in pairs], [b.cols for _, b in pairs], nnz)
        rng = np.random.default_rng(7)
        for it in range(4):
            if it:
                # New values in the same device arrays: numeric-only computes.
                new = []
                for (A, B), dA, dB in zip(pairs, As, Bs):
                    a2 = A.with_values(rng.uniform(-1, 1, A.nnz).astype(np.float32))
                    b2 = B.with_values(rng.uniform(-1, 1, B.nnz).astype(np.float32))
                    dA.set_values(a2.values)
                    dB.set_values(b2.values)
                    new.append((a2, b2))
                pairs = new
            F.check(bt.compute(As, Bs, out.structs), "batch_compute")
            assert_pairs_match(pairs, [out.download(p) for p in range(len(pairs))])
        # The first compute fills, the second (numeric-only) runs eagerly and
        # moves the groups to their position maps, the third is captured and
        # launched as a graph, later ones replay it.
        assert bt.graph_launches() == (2 if graph else 0)
    finally:
        if out is not None:
            out.free()
        bt.close()
        for d in owned:
            d.free()


@pytest.mark.parametrize("graph", [True, False])
def test_square_pairs_share_one_value_array(gpu, graph):
    """C = A*A for every pair through one device matrix per pair (its value
    array is both operands' values): the batch stages only A's values, and
    new values in that array reach both operands on every compute."""
    mats = [gen.uniform_k(n, 8, seed=s) for s, n in enumerate((64, 300, 1024, 2048, 513))]
    dmats = [rvsp.DeviceCsr.upload(M) for M in mats]
    bt = Batch(graph=graph)
    out = None
    try:
        nnz = bt.estimate(dmats, dmats)
        out = BatchOutput([M.rows for M in mats], [M.cols for M in mats], nnz)
        rng = np.random.default_rng(11)
        for it in range(4):
            if it:
                mats = [M.with_values(rng.uniform(0.5, 1.5, M.nnz).astype(np.float32))
                        for M in mats]
                for M, dM in zip(mats, dmats):
                    dM.set_values(M.values)
            F.check(bt.compute(dmats, dmats, out.structs), "batch_compute")
            assert_pairs_match([(M, M) for M in mats],
                               [out.download(p) for p in range(len(mats))])
        assert bt.graph_launches() == (2 if graph else 0)
    finally:
        if out is not None:
            out.free()
        bt.close()
        for d in dmats:
            d.free()


def test_batch_edge_pairs(gpu):
    """Pairs with empty rows, an all-empty product and 1x1 operands."""
    empty = Csr(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    B = Csr.from_dense(np.arange(63, dtype=np.float32).reshape(7, 9) % 3)
    A = Csr.from_dense(np.array([[0, 1, 0, 0, 0, 0, 2], [0] * 7, [3, 0, 0, 0, 0, 0, 0]],
                                np.float32))
    one = Csr.from_dense(np.array([[2.0]], np.float32))
    pairs = [(A, B), (empty, B), (one, one), (A, B)]
    got, _ = spgemm_batch(pairs)
    assert_pairs_match(pairs, got)
    assert got[1].nnz == 0


def test_batch_argument_checks(gpu):
    pairs = small_pairs(count=3, seed=3)
    As, Bs, owned = upload_pairs(pairs)
    try:
        a = Batch._structs(As)
        b = Batch._structs(Bs)
        out = (C.c_int32 * 3)()
        h = C.c_void_p()
        F.check(L.rvsp_spgemm_descr_create(C.byref(h)))
        # Host-memory descriptors have no batched path.
        assert L.rvsp_spgemm_batch_work_estimation(h, 3, a, b, out) == F.ERROR_INVALID_ARGUMENT
        F.check(L.rvsp_spgemm_set_memory(h, F.MEMORY_DEVICE))
        assert L.rvsp_spgemm_batch_compute(h, 3, a, b, (F.CsrMatrix * 3)()) == \
            F.ERROR_INVALID_ARGUMENT  # before estimation
        assert L.rvsp_spgemm_batch_work_estimation(h, 0, a, b, out) == F.ERROR_INVALID_ARGUMENT
        assert L.rvsp_spgemm_batch_work_estimation(h, 3, None, b, out) == F.ERROR_NULL_POINTER
        bad = Batch._structs(Bs)
        bad[1].rows += 1  # A[1].cols != B[1].rows
        assert L.rvsp_spgemm_batch_work_estimation(h, 3, a, bad, out) == F.ERROR_INVALID_ARGUMENT
        F.check(L.rvsp_spgemm_batch_work_estimation(h, 3, a, b, out))
        assert L.rvsp_spgemm_batch_groups(h) == 1
        nnz = np.array(out[:])
        o = BatchOutput([x.rows for x, _ in pairs], [y.cols for _, y in pairs], nnz)
        try:
            assert L.rvsp_spgemm_batch_compute(h, 2, a, b, o.structs) == F.ERROR_INVALID_ARGUMENT
            small = (F.CsrMatrix * 3)(*o.structs)
            small[2].nnz = int(nnz[2]) - 1
            if nnz[2] > 0:
                assert L.rvsp_spgemm_batch_compute(h, 3, a, b, small) == F.ERROR_INVALID_ARGUMENT
            F.check(L.rvsp_spgemm_batch_compute(h, 3, a, b, o.structs))
            assert_pairs_match(pairs, [o.download(p) for p in range(3)])
            # set_algo after a batch estimation is refused, as after one.
            assert L.rvsp_spgemm_set_algo(h, F.ALGO_HIP_FIXED) == F.ERROR_INVALID_ARGUMENT
        finally:
            o.free()
        L.rvsp_spgemm_descr_destroy(h)
    finally:
        for d in owned:
            d.free()


def test_batch_learned_and_tuned_algos_keep_bits(gpu):
    pairs = small_pairs(count=10, seed=5)
    for algo in (F.ALGO_HIP_TUNED, F.ALGO_HIP_LEARNED):
        got, _ = spgemm_batch(pairs, algo=algo)
        assert_pairs_match(pairs, got)


def test_full_size_2048_pairs_across_int32_groups(gpu):
    """The full batched workload: 2048 pairs, sizes log-uniform 256-65536,
    lognormal(ln 8, 1) entries per row capped at 256 (4.2e9 products, so the
    stack needs at least two int32 groups). Every output word of every pair
    equals the per-pair descriptor path's (order-sensitive checksums over
    each array), first compute and graph replays alike, and sampled pairs
    from both groups equal the oracle bit for bit."""
    import torch
    from rvsp import devgen
    pairs, keep = devgen.lognormal_batch(count=2048, seed=0, device="cuda")
    n = len(pairs)
    L.rvsp_hip_pool_release()

    def out_arrays(nnz):
        tot = int(sum(nnz))
        ci = torch.full((max(tot, 1),), -1, dtype=torch.int32, device="cuda")
        val = torch.full((max(tot, 1),), -1, dtype=torch.int32, device="cuda").view(torch.float32)
        rps = [torch.full((a.rows + 1,), -1, dtype=torch.int32, device="cuda") for a, _ in pairs]
        st = (F.CsrMatrix * n)()
        off = 0
        for p, (A, B) in enumerate(pairs):
            st[p] = F.CsrMatrix(A.rows, B.cols, int(nnz[p]), rps[p].data_ptr(),
                                ci.data_ptr() + 4 * off, val.data_ptr() + 4 * off,
                                F.DTYPE_FP32, F.FORMAT_CSR, 0)
            off += int(nnz[p])
        return ci, val, rps, st

    def checksum(t):
        w = t.reshape(-1).view(torch.int32).long()
        idx = torch.arange(w.numel(), device=w.device, dtype=torch.long) * 2 + 1
        return int((w * idx).sum()) % (1 << 64)

    # Per-pair descriptors (FIXED), each pair's arrays checksummed.
    want = []
    for A, B in pairs:
        d = rvsp.Descriptor(algo=F.ALGO_HIP_FIXED, memory=F.MEMORY_DEVICE)
        nnz = d.estimate(A, B)
        ci = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda")
        val = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda").view(torch.float32)
        rp = torch.full((A.rows + 1,), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        cm = F.CsrMatrix(A.rows, B.cols, nnz, rp.data_ptr(), ci.data_ptr(), val.data_ptr(),
                         F.DTYPE_FP32, F.FORMAT_CSR, 0)
        F.check(d.compute(A, B, cm))
        want.append((nnz, checksum(rp), checksum(ci[:nnz]), checksum(val[:nnz])))
        d.close()
    bt = Batch(algo=F.ALGO_HIP_LEARNED)
    As, Bs = [a for a, _ in pairs], [b for _, b in pairs]
    nnz_b = bt.estimate(As, Bs)
    assert bt.groups() >= 2
    assert [int(x) for x in nnz_b] == [w[0] for w in want]
    ci, val, rps, st = out_arrays(nnz_b)
    # torch fills on its own stream; the library's work runs on another
    # (non-blocking) stream, so the fills complete first.
    torch.cuda.synchronize()
    offs = np.concatenate([[0], np.cumsum(np.asarray(nnz_b, np.int64))])

    def check_all():
        for p in range(n):
            k = int(nnz_b[p])
            got = (checksum(rps[p]), checksum(ci[offs[p]:offs[p] + k]),
                   checksum(val[offs[p]:offs[p] + k]))
            assert got == want[p][1:], f"pair {p}"

    F.check(bt.compute(As, Bs, st))
    check_all()
    for _ in range(3):  # position maps, capture, replay
        F.check(bt.compute(As, Bs, st))
    assert bt.graph_launches() >= 1
    val.view(torch.int32).fill_(-1)
    torch.cuda.synchronize()
    F.check(bt.compute(As, Bs, st))  # a graph replay rewrites every value
    check_all()
    # Oracle on sampled pairs of both groups.
    for p in (0, 7, n // 2 - 1, n // 2, n - 8, n - 1):
        A, B = pairs[p][0].to_host(), pairs[p][1].to_host()
        rp_w, ci_w, v_w = oracle.spgemm(A, B)
        k = int(nnz_b[p])
        assert rps[p].cpu().numpy().tobytes() == rp_w.tobytes()
        assert ci[offs[p]:offs[p] + k].cpu().numpy().tobytes() == ci_w.tobytes()
        assert val[offs[p]:offs[p] + k].cpu().numpy().tobytes() == v_w.tobytes()
    bt.close()
    del ci, val, rps, keep
    torch.cuda.empty_cache()
    L.rvsp_hip_pool_release()
