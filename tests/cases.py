"""Named, seeded SpGEMM cases shared by the tests and the golden-vector script.

Each case builds (A, B) from rvsp.gen with explicit parameters; the golden
digests in tests/golden/digests.json were recorded from the reference
library on exactly these inputs.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rv-sparse_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

from rvsp import gen  # noqa: E402
from rvsp.api import Csr  # noqa: E402


def _pairs(count, n_max, seed):
    return gen.batch_pairs(count=count, n_min=256, n_max=n_max, seed=seed)


def _rect(seed):
    # Rectangular operands with empty rows in A and in B.
    rng = np.random.Generator(np.random.PCG64(seed))
    ca = rng.poisson(3.0, size=300)
    ca[::7] = 0
    rp, ci = gen.random_rows(300, 500, np.minimum(ca, 500), rng)
    A = Csr(300, 500, rp, ci, rng.uniform(-1, 1, ci.size))
    cb = rng.poisson(6.0, size=500)
    cb[::5] = 0
    rp, ci = gen.random_rows(500, 200, np.minimum(cb, 200), rng)
    B = Csr(500, 200, rp, ci, rng.uniform(-1, 1, ci.size))
    return A, B


def _laplacian(g, iteration):
    A = gen.laplacian_2d(g)
    if iteration is not None:
        A = A.with_values(gen.laplacian_values(A, iteration))
    return A, A


def _long_rows(seed):
    # A few A rows reaching thousands of B rows, so C rows run long
    # (block / bitmap / global kernels).
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 20000
    counts = rng.integers(1, 6, size=n)
    counts[::997] = 3000
    rp, ci = gen.random_rows(n, n, counts, rng)
    A = Csr(n, n, rp, ci, rng.uniform(-1, 1, ci.size))
    cb = rng.integers(0, 40, size=n)
    rp, ci = gen.random_rows(n, n, cb, rng)
    B = Csr(n, n, rp, ci, rng.uniform(-1, 1, ci.size))
    return A, B


def build(name):
    """(A, B) for the named case."""
    if name == "cfg1_uniform16_n4096":
        A = gen.uniform_k(4096, 16, seed=0)
        return A, A
    if name == "cfg2_laplacian_g1024_it0":
        return _laplacian(1024, 0)
    if name == "cfg2_laplacian_g64_stencil":
        return _laplacian(64, None)
    if name == "cfg3_rmat_s12":
        A = gen.rmat(12, seed=0)
        return A, A
    if name == "cfg3_rmat_s16_seed3":
        A = gen.rmat(16, seed=3)
        return A, A
    if name.startswith("cfg4_pair"):
        i = int(name.rsplit("_", 1)[1])
        return _pairs(4, 8192, 0)[i]
    if name == "cfg5_banded_small":
        return gen.banded_mixed(rows=20000, inner=8000, cols=20000, dense_nnz=500, seed=0)
    if name == "rect_empty_rows":
        return _rect(7)
    if name == "long_rows":
        return _long_rows(11)
    raise KeyError(name)


GOLDEN_CASES = [
    "cfg1_uniform16_n4096",
    "cfg2_laplacian_g1024_it0",
    "cfg2_laplacian_g64_stencil",
    "cfg3_rmat_s12",
    "cfg3_rmat_s16_seed3",
    "cfg4_pair_0", "cfg4_pair_1", "cfg4_pair_2", "cfg4_pair_3",
    "cfg5_banded_small",
    "rect_empty_rows",
    "long_rows",
]


def laplacian_squared_nnz(g):
    """Closed form: the square of the 5-point stencil is the 13-point stencil
    with offsets (0,0), (+-1,0), (0,+-1), (+-2,0), (0,+-2), (+-1,+-1); an
    offset (dx, dy) stays inside a g x g grid at (g-|dx|)(g-|dy|) points."""
    offsets = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (2, 0), (-2, 0), (0, 2), (0, -2),
               (1, 1), (1, -1), (-1, 1), (-1, -1)]
    return sum((g - abs(dx)) * (g - abs(dy)) for dx, dy in offsets)


# ---------------------------------------------------------------------------
# Canonical-CSR check cases: seeded valid structures (empty rows, rows longer
# than one workgroup step of the device check) and the same structures with
# planted faults. Every row pointer stays within [0, nnz], so the reference's
# host check never reads past the column array.

def _csr_base(rng, rows, cols):
    lens = rng.poisson(4.0, size=rows)
    lens[rng.random(rows) < 0.15] = 0
    long_rows = rng.choice(rows, size=max(1, rows // 60), replace=False)
    lens[long_rows] = rng.integers(900, 2600, size=long_rows.size)
    lens = np.minimum(lens, cols)
    rp, ci = gen.random_rows(rows, cols, lens, rng)
    return np.asarray(rp, np.int32), np.asarray(ci, np.int32)


def _plant(rng, rp, ci, cols, kind):
    rp, ci = rp.copy(), ci.copy()
    rows = rp.size - 1
    nonempty = np.flatnonzero(np.diff(rp) > 0)
    multi = np.flatnonzero(np.diff(rp) > 1)
    if kind == "swap" and multi.size:
        r = rng.choice(multi)
        p = rng.integers(rp[r], rp[r + 1] - 1)
        ci[p], ci[p + 1] = ci[p + 1], ci[p]
    elif kind == "dup" and multi.size:
        r = rng.choice(multi)
        p = rng.integers(rp[r] + 1, rp[r + 1])
        ci[p] = ci[p - 1]
    elif kind == "high" and nonempty.size:
        ci[rng.integers(0, ci.size)] = cols + int(rng.integers(0, 3))
    elif kind == "low" and nonempty.size:
        ci[rng.integers(0, ci.size)] = -1 - int(rng.integers(0, 3))
    elif kind == "descend":
        r = int(rng.integers(0, rows))
        if rp[r] > 0:
            rp[r + 1] = rp[r] - 1
    elif kind == "head":
        rp[0] = 1
    return rp, ci


def csr_check_cases(n_seeds=12):
    """(name, rows, cols, row_ptr, col_idx) tuples; col_idx may be None."""
    out = []
    kinds = ["swap", "dup", "high", "low", "descend"]
    for seed in range(n_seeds):
        rng = np.random.Generator(np.random.PCG64(1000 + seed))
        rows = int(rng.integers(50, 400))
        cols = int(rng.integers(2700, 6000))
        rp, ci = _csr_base(rng, rows, cols)
        out.append((f"valid_{seed}", rows, cols, rp, ci))
        for k in kinds:
            out.append((f"{k}_{seed}", rows, cols) + _plant(rng, rp, ci, cols, k))
        # Two or three faults at once: the first in scan order must win.
        r2, c2 = rp, ci
        for k in rng.choice(kinds, size=int(rng.integers(2, 4))):
            r2, c2 = _plant(rng, r2, c2, cols, str(k))
        out.append((f"multi_{seed}", rows, cols, r2, c2))
    rng = np.random.Generator(np.random.PCG64(99))
    rp, ci = _csr_base(rng, 120, 3000)
    out.append(("head", 120, 3000) + _plant(rng, rp, ci, 3000, "head"))
    out.append(("no_rows", 0, 5, np.zeros(1, np.int32), np.zeros(0, np.int32)))
    out.append(("all_empty", 7, 5, np.zeros(8, np.int32), np.zeros(0, np.int32)))
    out.append(("no_cols_empty", 4, 5, np.zeros(5, np.int32), None))
    out.append(("no_cols_nonempty", 4, 5, np.array([0, 0, 2, 2, 3], np.int32), None))
    out.append(("no_cols_descend_first", 4, 5, np.array([0, 0, 0, -1, 0], np.int32), None))
    out.append(("one_entry", 1, 1, np.array([0, 1], np.int32), np.array([0], np.int32)))
    out.append(("first_col_negative", 2, 4, np.array([0, 1, 2], np.int32),
                np.array([-1, 0], np.int32)))
    return out


# ---------------------------------------------------------------------------
# Legacy-dtype products (FP64 -> FP64, INT8 -> INT32): seeded canonical
# operands. Values come from small integer sets so that many sums cancel to
# exactly zero (those entries are dropped); FP64 cases also carry fractional
# values, where the order of the additions shows in the last bits.

def legacy_pair(seed, i8, rows=120, inner=90, cols=3000, mean=4.0, long_rows=2):
    rng = np.random.Generator(np.random.PCG64(5000 + seed))
    la = np.minimum(rng.poisson(mean, size=rows), inner)
    la[rng.choice(rows, size=min(long_rows, rows), replace=False)] = inner  # wide C rows
    lb = np.minimum(rng.poisson(mean * 4, size=inner), cols)
    lb[rng.random(inner) < 0.1] = 0
    rpa, cia = gen.random_rows(rows, inner, la, rng)
    rpb, cib = gen.random_rows(inner, cols, lb, rng)
    if i8:
        va = rng.integers(-128, 128, size=len(cia)).astype(np.int8)
        vb = rng.choice(np.array([-2, -1, 1, 2], np.int8), size=len(cib))
        dt = np.int8
    else:
        va = rng.integers(-2, 3, size=len(cia)).astype(np.float64)
        va[va == 0] = 1.0
        vb = rng.choice(np.array([-1.0, 1.0, 0.5, -0.5]), size=len(cib))
        frac = rng.random(len(cib)) < 0.3
        vb[frac] = rng.standard_normal(int(frac.sum()))
        dt = np.float64
    return (Csr(rows, inner, rpa, cia, va, dtype=dt), Csr(inner, cols, rpb, cib, vb, dtype=dt))
