"""Batched products over device-resident pairs (rvsp_spgemm_batch_* in
include/rv_sparse_hip.h)."""
import ctypes as C

import numpy as np

from . import _ffi as F
from ._ffi import check, lib
from .api import DeviceCsr, device_alloc, device_free, from_device


class Batch:
    """A descriptor used for batched products C[p] = A[p] * B[p]."""

    def __init__(self, algo=F.ALGO_DEFAULT, group_limit=None, graph=True, choices=None,
                 fuse=False, persistent=1):
        self.handle = C.c_void_p()
        check(lib().rvsp_spgemm_descr_create(C.byref(self.handle)), "descr_create")
        check(lib().rvsp_spgemm_set_algo(self.handle, algo), "set_algo")
        check(lib().rvsp_spgemm_set_memory(self.handle, F.MEMORY_DEVICE), "set_memory")
        if choices:
            arr = (F.BinChoice * len(choices))(*[F.BinChoice(*c) for c in choices])
            check(lib().rvsp_spgemm_set_bin_choices(self.handle, arr, len(choices)),
                  "set_bin_choices")
        if group_limit is not None:
            check(lib().rvsp_spgemm_batch_set_group_limit(self.handle, group_limit),
                  "batch_set_group_limit")
        check(lib().rvsp_spgemm_batch_set_graph(self.handle, 1 if graph else 0), "set_graph")
        check(lib().rvsp_spgemm_batch_set_fuse(self.handle, 1 if fuse else 0), "set_fuse")
        # persistent: 0 per-bin launches, 1 measured choice, 2 always the
        # persistent launch (True/False accepted for 2/0).
        mode = {True: 2, False: 0}.get(persistent, persistent) if isinstance(persistent, bool) \
            else int(persistent)
        check(lib().rvsp_spgemm_batch_set_persistent(self.handle, mode), "set_persistent")
        self.c_nnz = None

    @staticmethod
    def _structs(mats):
        """A ctypes CsrMatrix array for the operands (returned as is when it
        already is one, so callers can build it once for repeated computes)."""
        if isinstance(mats, C.Array):
            return mats
        arr = (F.CsrMatrix * len(mats))()
        for i, m in enumerate(mats):
            arr[i] = m if isinstance(m, F.CsrMatrix) else m.struct()
        return arr

    def estimate(self, As, Bs):
        a, b = self._structs(As), self._structs(Bs)
        out = (C.c_int32 * len(As))()
        check(lib().rvsp_spgemm_batch_work_estimation(self.handle, len(As), a, b, out),
              "batch_work_estimation")
        self.c_nnz = np.array(out[:], dtype=np.int64)
        return self.c_nnz

    def compute(self, As, Bs, Cs):
        """Cs: a ctypes array of CsrMatrix (updated in place); returns status."""
        a, b = self._structs(As), self._structs(Bs)
        return lib().rvsp_spgemm_batch_compute(self.handle, len(As), a, b, Cs)

    def groups(self):
        return int(lib().rvsp_spgemm_batch_groups(self.handle))

    def graph_launches(self):
        return int(lib().rvsp_spgemm_batch_graph_launches(self.handle))

    def stream_segments(self):
        return int(lib().rvsp_spgemm_batch_stream_segments(self.handle))

    def numeric_path(self):
        """(path, persistent_ms, per_bin_ms): path 1 persistent, 2 per bin."""
        a, b = C.c_float(), C.c_float()
        path = int(lib().rvsp_spgemm_batch_numeric_path(self.handle, C.byref(a), C.byref(b)))
        return path, a.value, b.value

    def stream(self):
        s = C.c_void_p()
        check(lib().rvsp_spgemm_get_stream(self.handle, C.byref(s)), "get_stream")
        return s.value

    def set_async(self, on=True):
        check(lib().rvsp_spgemm_set_async(self.handle, 1 if on else 0), "set_async")

    def close(self):
        if self.handle:
            lib().rvsp_spgemm_descr_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchOutput:
    """Device C arrays for a batch: one allocation whose per-pair views lie
    back to back (written in place), or one allocation per pair."""

    def __init__(self, rows, cols, c_nnz, contiguous=True):
        self.n = len(rows)
        self.rows, self.cols = list(rows), list(cols)
        self.c_nnz = [int(x) for x in c_nnz]
        self.bufs = []
        self.structs = (F.CsrMatrix * self.n)()
        off = np.concatenate([[0], np.cumsum(self.c_nnz)]).astype(np.int64)
        self.rp = [device_alloc(4 * (r + 1)) for r in self.rows]
        self.bufs += self.rp
        if contiguous:
            total = max(int(off[-1]), 1)
            ci, val = device_alloc(4 * total), device_alloc(4 * total)
            self.bufs += [ci, val]
            self.ci = [ci + 4 * int(off[p]) for p in range(self.n)]
            self.val = [val + 4 * int(off[p]) for p in range(self.n)]
        else:
            self.ci = [device_alloc(4 * max(k, 1)) for k in self.c_nnz]
            self.val = [device_alloc(4 * max(k, 1)) for k in self.c_nnz]
            self.bufs += self.ci + self.val
        for p in range(self.n):
            self.structs[p] = F.CsrMatrix(self.rows[p], self.cols[p], self.c_nnz[p], self.rp[p],
                                          self.ci[p], self.val[p], F.DTYPE_FP32, F.FORMAT_CSR, 0)

    def download(self, p):
        from .api import Csr
        k = self.c_nnz[p]
        return Csr(self.rows[p], self.cols[p], from_device(self.rp[p], self.rows[p] + 1, np.int32),
                   from_device(self.ci[p], k, np.int32), from_device(self.val[p], k, np.float32))

    def free(self):
        for b in self.bufs:
            device_free(b)
        self.bufs = []


def upload_pairs(pairs):
    """DeviceCsr operands for host pairs; a host matrix object used several
    times is uploaded once (so the library stacks it once)."""
    seen = {}

    def up(m):
        if id(m) not in seen:
            seen[id(m)] = DeviceCsr.upload(m)
        return seen[id(m)]

    As = [up(a) for a, _ in pairs]
    Bs = [up(b) for _, b in pairs]
    return As, Bs, list(seen.values())


def spgemm_batch(pairs, algo=F.ALGO_DEFAULT, contiguous=True, group_limit=None, graph=True,
                 repeats=1, choices=None, fuse=False, persistent=1):
    """Host pairs in, host products out, through the batched device path."""
    As, Bs, owned = upload_pairs(pairs)
    bt = Batch(algo=algo, group_limit=group_limit, graph=graph, choices=choices, fuse=fuse,
               persistent=persistent)
    out = None
    try:
        nnz = bt.estimate(As, Bs)
        out = BatchOutput([a.rows for a, _ in pairs], [b.cols for _, b in pairs], nnz,
                          contiguous=contiguous)
        for _ in range(repeats):
            check(bt.compute(As, Bs, out.structs), "batch_compute")
        return [out.download(p) for p in range(len(pairs))], bt.groups()
    finally:
        if out is not None:
            out.free()
        bt.close()
        for d in owned:
            d.free()
