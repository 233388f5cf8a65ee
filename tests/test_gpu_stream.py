"""Persistent row-stream numeric pass of batched products: numeric-only
batch computes run every group's short-row numeric bins as one launch over a
product-balanced work queue. Every pair must stay bit-identical to the
oracle, with the pass on and off, eager and replayed from the graph, with new
values each time and outputs pre-filled with a sentinel, including rows whose
products collide on the same C entry many times inside one 64-product window
(order rounds, and the lane-order fallback past 7 repeats)."""
import numpy as np
import pytest

from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen
from rvsp.api import Csr, to_device
from rvsp.batch import Batch, BatchOutput, upload_pairs

pytestmark = pytest.mark.gpu


def collision_pairs():
    rng = np.random.default_rng(11)
    pairs = []
    # Every B row holds column 0 and two of columns 1..4: a C row gets one
    # product per A entry on column 0 (dozens inside one window) and
    # repeats on the others.
    n = 96
    b = np.zeros((n, 40), np.float32)
    for k in range(n):
        b[k, 0] = rng.uniform(-1, 1)
        b[k, 1 + rng.choice(4, 2, replace=False)] = rng.uniform(-1, 1, 2)
    B = Csr.from_dense(b)
    mask = rng.random((300, n)) < rng.uniform(0.05, 0.9, (300, 1))
    A = Csr.from_dense((mask * rng.uniform(-1, 1, (300, n))).astype(np.float32))
    pairs.append((A, B))
    # C rows of about 600 to 2600 positions (the stream's long rows, walked
    # once per range of accumulators).
    A2 = gen.lognormal_rows(3000, mu=np.log(150.0), sigma=0.5, cap=400, seed=12)
    B2 = gen.lognormal_rows(3000, mu=np.log(8.0), sigma=0.5, cap=16, seed=13)
    pairs.append((A2, B2))
    # Mostly short rows and a tail of rows with up to about 10000 products
    # and 5000 positions: past 4096 products, rows left to the long-row
    # kernels, next to the stream's.
    A4 = gen.lognormal_rows(6000, mu=np.log(20.0), sigma=1.2, cap=500, seed=17)
    B4 = gen.lognormal_rows(6000, mu=np.log(20.0), sigma=0.3, cap=40, seed=18)
    pairs.append((A4, B4))
    # Rows fed by one B row (identity positions) next to mixed rows.
    A3 = gen.uniform_k(1500, 1, seed=14)
    B3 = gen.uniform_k(1500, 30, seed=15)
    pairs.append((A3, B3))
    pairs += gen.batch_pairs(12, n_min=64, n_max=4096, seed=16)
    return pairs


def assert_pairs_match(pairs, got):
    for p, ((A, B), g) in enumerate(zip(pairs, got)):
        want = oracle.spgemm(A, B)
        for name, x, w in zip(("row_ptr", "col_idx", "values"),
                              (g.row_ptr, g.col_idx, g.values), want):
            assert x.tobytes() == w.tobytes(), f"pair {p} {name}"


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("algo", [F.ALGO_HIP_FIXED, F.ALGO_HIP_LEARNED])
def test_stream_pass_matches_oracle_over_repeats(gpu, graph, algo):
    pairs = collision_pairs()
    As, Bs, owned = upload_pairs(pairs)
    bt = Batch(algo=algo, graph=graph, group_limit=400_000, persistent=True)
    out = None
    rng = np.random.default_rng(21)
    try:
        nnz = bt.estimate(As, Bs)
        assert bt.groups() >= 2
        out = BatchOutput([a.rows for a, _ in pairs], [b.cols for _, b in pairs], nnz)
        for it in range(5):
            if it:
                new = []
                for (A, B), dA, dB in zip(pairs, As, Bs):
                    a2 = A.with_values(rng.uniform(-1, 1, A.nnz).astype(np.float32))
                    b2 = B.with_values(rng.uniform(-1, 1, B.nnz).astype(np.float32))
                    dA.set_values(a2.values)
                    dB.set_values(b2.values)
                    new.append((a2, b2))
                pairs = new
                for p in range(len(pairs)):  # every value must be rewritten
                    to_device(out.val[p], np.full(out.c_nnz[p], -1, np.int32))
            F.check(bt.compute(As, Bs, out.structs), "batch_compute")
            assert_pairs_match(pairs, [out.download(p) for p in range(len(pairs))])
            if it >= 1:
                assert bt.stream_segments() > 0
        assert bt.graph_launches() == (3 if graph else 0)
    finally:
        if out is not None:
            out.free()
        bt.close()
        for d in owned:
            d.free()


def test_stream_pass_on_and_off_agree_on_uneven_pairs(gpu):
    pairs = gen.batch_pairs(40, n_min=32, n_max=8192, seed=31)
    res = {}
    for persistent in (True, False):
        As, Bs, owned = upload_pairs(pairs)
        bt = Batch(algo=F.ALGO_HIP_LEARNED, persistent=persistent)
        out = None
        try:
            nnz = bt.estimate(As, Bs)
            out = BatchOutput([a.rows for a, _ in pairs], [b.cols for _, b in pairs], nnz,
                              contiguous=False)
            for _ in range(4):
                F.check(bt.compute(As, Bs, out.structs), "batch_compute")
            assert (bt.stream_segments() > 0) == persistent
            res[persistent] = [out.download(p) for p in range(len(pairs))]
        finally:
            if out is not None:
                out.free()
            bt.close()
            for d in owned:
                d.free()
    for g, h in zip(res[True], res[False]):
        assert g.values.tobytes() == h.values.tobytes()
        assert g.col_idx.tobytes() == h.col_idx.tobytes()
    assert_pairs_match(pairs, res[True])


def test_measured_choice_between_persistent_and_per_bin_launches(gpu):
    """Default mode: the first numeric-only compute times the persistent pass
    against the groups' per-bin launches and keeps the faster; the bits are
    the oracle's before, during and after the choice."""
    pairs = gen.batch_pairs(30, n_min=64, n_max=4096, seed=41)
    As, Bs, owned = upload_pairs(pairs)
    bt = Batch(algo=F.ALGO_HIP_LEARNED, group_limit=600_000)
    out = None
    try:
        nnz = bt.estimate(As, Bs)
        assert bt.groups() >= 2
        out = BatchOutput([a.rows for a, _ in pairs], [b.cols for _, b in pairs], nnz)
        F.check(bt.compute(As, Bs, out.structs), "batch_compute")
        assert bt.numeric_path()[0] == 1  # the first compute's search pass
        assert_pairs_match(pairs, [out.download(p) for p in range(len(pairs))])
        for _ in range(3):
            for p in range(len(pairs)):
                to_device(out.val[p], np.full(out.c_nnz[p], -1, np.int32))
            F.check(bt.compute(As, Bs, out.structs), "batch_compute")
            assert_pairs_match(pairs, [out.download(p) for p in range(len(pairs))])
        path, t_persistent, t_bin = bt.numeric_path()
        assert path in (1, 2)
        assert t_persistent > 0 and t_bin > 0
        assert (path == 1) == (t_persistent <= t_bin)
        assert (bt.stream_segments() > 0) == (path == 1)
    finally:
        if out is not None:
            out.free()
        bt.close()
        for d in owned:
            d.free()
