"""Seeded synthetic workloads for the SpGEMM benchmark configurations.

Every generator returns canonical CSR (ascending, distinct columns per row)
built with numpy's PCG64 generator from an explicit seed, so a workload is
fully named by its parameters:

  uniform_k      n x n, exactly k distinct uniform columns per row
  laplacian_2d   5-point Laplacian of a g x g grid (diag 4, off-diag -1)
  rmat           R-MAT graph adjacency, duplicates merged
  lognormal_rows n x n, per-row counts from a clamped lognormal
  banded_mixed   rectangular band plus 1% dense-ish rows
"""
import numpy as np

from .api import Csr


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def random_rows(rows, cols, counts, rng):
    """CSR structure with counts[i] distinct uniform columns in row i."""
    counts = np.asarray(counts, dtype=np.int64)
    if np.any(counts > cols):
        raise ValueError("row count exceeds the number of columns")
    row_ptr = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col = np.empty(int(row_ptr[-1]), dtype=np.int64)
    dense = counts * 4 > cols
    # Rows that ask for a large share of the columns: exact sampling without
    # replacement, one row at a time.
    for i in np.nonzero(dense)[0]:
        s, e = row_ptr[i], row_ptr[i + 1]
        col[s:e] = np.sort(rng.choice(cols, size=e - s, replace=False))
    # Other rows: draw with replacement, redraw duplicates until none remain.
    row_of = np.repeat(np.arange(rows, dtype=np.int64), counts)
    idx = np.nonzero(~dense[row_of])[0]
    if idx.size:
        row_of = row_of[idx]
        c = rng.integers(0, cols, size=idx.size, dtype=np.int64)
        while True:
            key = row_of * cols + c
            order = np.argsort(key, kind="stable")
            ks = key[order]
            dup = np.zeros(ks.size, dtype=bool)
            dup[1:] = ks[1:] == ks[:-1]
            if not dup.any():
                break
            redo = order[dup]
            c[redo] = rng.integers(0, cols, size=redo.size, dtype=np.int64)
        # idx is row-major, so the (row, col)-sorted draws line up with it.
        col[idx] = c[np.lexsort((c, row_of))]
    return row_ptr.astype(np.int32), col.astype(np.int32)


def uniform_k(n=4096, k=16, seed=0, low=-1.0, high=1.0):
    """n x n, exactly k distinct uniform columns per row, values U[low, high)."""
    rng = _rng(seed)
    rp, ci = random_rows(n, n, np.full(n, k), rng)
    vals = rng.uniform(low, high, size=ci.size).astype(np.float32)
    return Csr(n, n, rp, ci, vals)


def laplacian_2d(g=1024):
    """5-point Laplacian on a g x g grid: row r*g+c couples to its N/W/E/S
    neighbours with -1 and to itself with 4."""
    n = g * g
    r, c = np.divmod(np.arange(n, dtype=np.int64), g)
    offs = [(-g, r > 0), (-1, c > 0), (0, np.ones(n, bool)), (1, c < g - 1), (g, r < g - 1)]
    cols = np.stack([np.arange(n) + o for o, _ in offs], axis=1)
    mask = np.stack([m for _, m in offs], axis=1)
    vals = np.where(np.arange(5) == 2, 4.0, -1.0)[None, :].repeat(n, axis=0)
    counts = mask.sum(axis=1)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    return Csr(n, n, rp, cols[mask], vals[mask])


def laplacian_values(A, iteration):
    """Values redrawn U[0.5, 1.5) for numeric-only repeat `iteration`."""
    return _rng(iteration).uniform(0.5, 1.5, size=A.nnz).astype(np.float32)


def rmat(scale=20, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
    """R-MAT adjacency, 2^scale vertices, edge_factor * 2^scale edges drawn,
    duplicates merged, values U[0, 1) drawn after merging."""
    rng = _rng(seed)
    n = 1 << scale
    m = edge_factor * n
    row = np.zeros(m, dtype=np.int64)
    col = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    chunk = 1 << 22
    for s in range(0, m, chunk):
        e = min(m, s + chunk)
        rr = np.zeros(e - s, dtype=np.int64)
        cc = np.zeros(e - s, dtype=np.int64)
        for _ in range(scale):
            u = rng.random(e - s)
            right = ((u >= a) & (u < ab)) | (u >= abc)
            down = u >= ab
            rr = (rr << 1) | down
            cc = (cc << 1) | right
        row[s:e] = rr
        col[s:e] = cc
    key = np.unique(row * n + col)
    row, col = np.divmod(key, n)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(row, minlength=n), out=rp[1:])
    vals = rng.random(col.size).astype(np.float32)
    return Csr(n, n, rp, col, vals)


def lognormal_rows(n, mu=np.log(8.0), sigma=1.0, cap=256, rng=None, seed=0,
                   low=-1.0, high=1.0):
    """n x n, row counts clamp(round(lognormal(mu, sigma)), 1, min(cap, n))."""
    rng = rng if rng is not None else _rng(seed)
    counts = np.clip(np.rint(rng.lognormal(mu, sigma, size=n)), 1, min(cap, n)).astype(np.int64)
    rp, ci = random_rows(n, n, counts, rng)
    vals = rng.uniform(low, high, size=ci.size).astype(np.float32)
    return Csr(n, n, rp, ci, vals)


def batch_pairs(count=2048, n_min=256, n_max=65536, seed=0, **kw):
    """`count` independent (A, B) pairs, square sizes log-uniform in
    [n_min, n_max], lognormal row counts."""
    rng = _rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=count))).astype(np.int64)
    pairs = []
    for n in sizes:
        A = lognormal_rows(int(n), rng=rng, **kw)
        B = lognormal_rows(int(n), rng=rng, **kw)
        pairs.append((A, B))
    return pairs


def banded_mixed(rows=500_000, inner=200_000, cols=500_000, band=32, band_nnz=24,
                 dense_every=100, dense_nnz=2000, b_nnz=8, seed=0):
    """A (rows x inner): row i has band_nnz distinct columns inside a width-band
    window centred at floor(i * inner / rows), except every dense_every-th row
    (i % dense_every == 0), which has dense_nnz uniform columns. B (inner x
    cols): b_nnz uniform columns per row. Values N(0, 1)."""
    rng = _rng(seed)
    dense = (np.arange(rows) % dense_every) == 0
    counts = np.where(dense, dense_nnz, band_nnz).astype(np.int64)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    ci = np.empty(int(rp[-1]), dtype=np.int64)
    # Band rows: band_nnz distinct offsets out of `band` per row.
    nb = int((~dense).sum())
    offs = np.argsort(rng.random((nb, band)), axis=1)[:, :band_nnz]
    offs.sort(axis=1)
    centre = (np.nonzero(~dense)[0] * inner) // rows
    start = np.clip(centre - band // 2, 0, inner - band)
    band_cols = start[:, None] + offs
    idx = (rp[np.nonzero(~dense)[0]][:, None] + np.arange(band_nnz)[None, :]).ravel()
    ci[idx] = band_cols.ravel()
    drows = np.nonzero(dense)[0]
    _, d_ci = random_rows(drows.size, inner, np.full(drows.size, dense_nnz), rng)
    ci[(rp[drows][:, None] + np.arange(dense_nnz)[None, :]).ravel()] = d_ci
    A = Csr(rows, inner, rp, ci, rng.standard_normal(ci.size).astype(np.float32))
    b_rp, b_ci = random_rows(inner, cols, np.full(inner, b_nnz), rng)
    B = Csr(inner, cols, b_rp, b_ci, rng.standard_normal(b_ci.size).astype(np.float32))
    return A, B
