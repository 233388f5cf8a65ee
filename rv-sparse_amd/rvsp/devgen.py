"""Workload generators that build operands directly in device memory with
torch (used by bench.py for the full-size batched workload, where drawing
hundreds of millions of entries on the host would dominate the run).

The library itself never imports this module. Matrices come back as
`TorchCsr` views: plain device pointers and sizes, the form the C API takes.
"""
import math

import numpy as np

from . import _ffi as F


class TorchCsr:
    """A CSR matrix whose arrays are slices of torch device tensors."""

    def __init__(self, rows, cols, rp, ci, val):
        self.rows, self.cols = int(rows), int(cols)
        self.rp, self.ci, self.val = rp, ci, val  # int32, int32, float32 tensors
        self.nnz = int(ci.numel())

    def struct(self, nnz=None):
        return F.CsrMatrix(self.rows, self.cols, self.nnz if nnz is None else nnz,
                           self.rp.data_ptr(), self.ci.data_ptr(), self.val.data_ptr(),
                           F.DTYPE_FP32, F.FORMAT_CSR, 0)

    def to_host(self):
        from .api import Csr
        return Csr(self.rows, self.cols, self.rp.cpu().numpy(), self.ci.cpu().numpy(),
                   self.val.cpu().numpy())


def _distinct_columns(row_of, bound, gen, device):
    """Uniform columns in [0, bound[row]) per entry, redrawn until no row
    holds a column twice; returns the columns sorted within each row (entries
    are grouped by row in ascending order)."""
    import torch
    u = torch.rand(row_of.numel(), generator=gen, device=device, dtype=torch.float64)
    b = bound[row_of]
    col = torch.minimum((u * b).long(), b - 1)
    shift = 1 << int(math.ceil(math.log2(max(int(bound.max()), 2))))
    while True:
        key = row_of * shift + col
        key, order = torch.sort(key)
        dup = torch.zeros_like(key, dtype=torch.bool)
        dup[1:] = key[1:] == key[:-1]
        n_dup = int(dup.sum())
        col = key % shift  # sorted by (row, column): rows stay in place
        if n_dup == 0:
            return col
        u = torch.rand(n_dup, generator=gen, device=device, dtype=torch.float64)
        bd = b[dup]
        col[dup] = torch.minimum((u * bd).long(), bd - 1)


def lognormal_batch(count=2048, n_min=256, n_max=65536, mu=math.log(8.0), sigma=1.0,
                    cap=256, seed=0, device="cuda"):
    """`count` independent square pairs (A, B): sizes n = floor(exp(U[ln n_min,
    ln n_max])), row counts clamp(round(lognormal(mu, sigma)), 1, min(cap, n)),
    distinct uniform columns, values U[-1, 1). Returns (pairs, tensors)."""
    import torch
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=count))).astype(np.int64)
    n_mat = np.repeat(sizes, 2)  # A then B of every pair
    gen = torch.Generator(device=device).manual_seed(seed)
    n_t = torch.as_tensor(n_mat, device=device)
    row_bound = torch.repeat_interleave(n_t, n_t)  # columns of the row's matrix
    counts = torch.empty(row_bound.numel(), device=device, dtype=torch.float64)
    counts.log_normal_(mean=mu, std=sigma, generator=gen)
    counts = torch.clamp(torch.round(counts), min=1).long()
    counts = torch.minimum(counts, torch.clamp(row_bound, max=cap))
    row_of = torch.repeat_interleave(torch.arange(row_bound.numel(), device=device), counts)
    col = _distinct_columns(row_of, row_bound, gen, device).int()
    del row_of
    val = (torch.rand(col.numel(), generator=gen, device=device) * 2 - 1).float()
    # Row pointers: every matrix's own, starting at 0, back to back.
    ends = torch.cumsum(counts, 0)
    starts_row = np.concatenate([[0], np.cumsum(n_mat)])
    nm = len(n_mat)
    nnz_start = torch.zeros(nm + 1, dtype=torch.long, device=device)
    nnz_start[1:] = ends[torch.as_tensor(starts_row[1:] - 1, device=device)]
    m_of_row = torch.repeat_interleave(torch.arange(nm, device=device), n_t)
    rp_all = torch.empty(int(starts_row[-1]) + nm, dtype=torch.int32, device=device)
    pos = torch.arange(row_bound.numel(), device=device) + m_of_row + 1
    rp_all[pos] = (ends - nnz_start[m_of_row]).int()
    rp_all[torch.as_tensor(starts_row[:-1] + np.arange(nm), device=device)] = 0
    del m_of_row, pos, ends
    nnz_start_h = nnz_start.cpu().numpy()
    mats = []
    for m, n in enumerate(n_mat):
        b0 = int(starts_row[m]) + m
        s, e = int(nnz_start_h[m]), int(nnz_start_h[m + 1])
        mats.append(TorchCsr(n, n, rp_all[b0:b0 + n + 1], col[s:e], val[s:e]))
    pairs = [(mats[2 * p], mats[2 * p + 1]) for p in range(count)]
    return pairs, (rp_all, col, val)
