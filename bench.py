#!/usr/bin/env python3
"""SpGEMM throughput on one MI355X: GFLOP/s and HBM GB/s.

Workload (default): C = A * A for the 5-point Laplacian of a 1024 x 1024
grid (1,048,576 rows, FP32), structure analysed once by the descriptor, then
numeric-only multiplies with A's values redrawn U[0.5, 1.5) per multiply
(seed = multiply index). All value sets are staged in HBM before the timed
region; a step is one numeric-only multiply through rvsp_spgemm_compute on
device-resident operands. GFLOP/s = 2 * intermediate products / time, the
reference's definition (bench/bench.c:753-754).

Two plans are timed on the same staged inputs: the plain fixed per-bin
strategy (RVSP_SPGEMM_ALGO_HIP_FIXED) and the autotuned one
(RVSP_SPGEMM_ALGO_HIP_TUNED after rvsp_spgemm_autotune, run untimed during
set-up). `value` is the tuned plan; both outputs are checked bit for bit.

Prints one JSON line. `roofline` covers the numeric kernel: algorithmic
bytes per multiply (4(M+1) + 8 nnz(A) + 4(M+1) + 8 nnz(C), A counted once)
over its average duration from HIP events on the descriptor's stream.
`cpu_baseline` times the reference's own descriptor path (built from its
sources into oracle/_ref) on the host, one thread, on a bounded number of
the same multiplies; the project's oracle restatement stands in where the
reference library is absent. Its `all_cores` field gives the reference on up
to 16 host threads (row blocks balanced by products, one descriptor each).
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rv-sparse_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


class Hip:
    """The few HIP runtime calls the timing needs, from the runtime the
    library already loaded."""

    def __init__(self):
        h = C.CDLL("libamdhip64.so.7")
        P = C.c_void_p
        h.hipEventCreate.argtypes = [C.POINTER(P)]
        h.hipEventRecord.argtypes = [P, P]
        h.hipEventSynchronize.argtypes = [P]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), P, P]
        h.hipEventDestroy.argtypes = [P]
        self.h = h

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        assert self.h.hipEventSynchronize(b) == 0
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def cpu_baseline(A, values, budget_s):
    from oracle import oracle
    from rvsp import _ffi as F
    if oracle.ref_available():
        probe, _ = oracle.time_reference_numeric(A, A, 1, F.CsrMatrix, values[:1])
        reps = max(3, min(len(values), int(budget_s / max(probe[0], 1e-6))))
        times, _ = oracle.time_reference_numeric(A, A, reps, F.CsrMatrix, values)
        kind = "reference"
    else:
        probe, _ = oracle.time_port_numeric(A, A, 1)
        reps = max(3, int(budget_s / max(probe[0], 1e-6)))
        times, _ = oracle.time_port_numeric(A, A, reps)
        kind = "port"
    return statistics.median(times), reps, kind


def run_plan(hip, A, dA, val_bufs, algo, steps, warmup, autotune_reps):
    """Analyse, (autotune), warm up, then time `steps` numeric-only multiplies.
    Returns timings, the plan and the device C (caller frees)."""
    import rvsp
    from rvsp import _ffi as F
    d = rvsp.Descriptor(algo=algo, memory=F.MEMORY_DEVICE)
    hip.sync()
    t0 = time.perf_counter()
    nnz_c = d.estimate(dA, dA)
    t_sym = time.perf_counter() - t0
    dC = rvsp.DeviceCsr(A.rows, A.cols, nnz_c)
    Cm = dC.struct(nnz_c)
    a = dA.struct()
    ws = C.c_void_p(d.ws.ctypes.data)
    L = F.lib()
    t0 = time.perf_counter()
    F.check(L.rvsp_spgemm_compute(d.handle, C.byref(a), C.byref(a), C.byref(Cm), ws),
            "first compute")
    t_first = time.perf_counter() - t0
    t_tune = None
    if autotune_reps:
        t0 = time.perf_counter()
        F.check(L.rvsp_spgemm_autotune(d.handle, C.byref(a), C.byref(a), C.byref(Cm), ws,
                                       autotune_reps), "autotune")
        t_tune = time.perf_counter() - t0
    d.set_async(True)
    stream = d.stream()

    def step(i):
        a.values = val_bufs[i]  # A and B are the same matrix; only values change
        return L.rvsp_spgemm_compute(d.handle, C.byref(a), C.byref(a), C.byref(Cm), ws)

    for i in range(warmup):
        F.check(step(steps + i), "warmup")
    # One event pair on the descriptor's stream brackets the whole timed
    # region: its average per multiply is the launches' device time (the
    # kernel plus the gap to the next launch). Events recorded around every
    # step would add their own packets between the multiplies.
    ev0, ev1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(ev0, stream)
    for i in range(steps):
        st = step(i)
        if st != 0:
            F.check(st, "compute")
    hip.record(ev1, stream)
    hip.sync()
    wall = time.perf_counter() - t0
    kernel_ms = hip.elapsed_ms(ev0, ev1) / steps
    plan = [p for p in d.plan() if p["phase"] == F.PHASE_NUMERIC and p["rows"] > 0]
    d.close()
    return {"ms_per_step": wall * 1e3 / steps, "kernel_ms": kernel_ms, "t_sym": t_sym,
            "t_first": t_first, "t_tune": t_tune, "plan": plan, "nnz_c": nnz_c, "C": dC}


# ---------------------------------------------------------------------------
# Secondary workloads: reported under "configs" next to the headline.

def cpu_oneshot(pairs, budget_s):
    """One-shot CPU products (reference raw scalar kernel when built, else
    the oracle restatement) over a prefix of `pairs` that fits the budget.
    Returns (seconds, products, n_pairs, kind)."""
    from oracle import oracle
    kind = "reference" if oracle.ref_available() else "port"
    spent, products, n = 0.0, 0, 0
    for A, B in pairs:
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_spgemm(A, B)
        else:
            oracle.spgemm(A, B)
        spent += time.perf_counter() - t0
        products += int(np.diff(B.row_ptr)[A.col_idx].sum())
        n += 1
        if spent > budget_s:
            break
    return spent, products, n, kind


def word_checksum(t, offset=0, chunk=1 << 27):
    """Order-sensitive checksum (mod 2^64) of a 32-bit device tensor's words,
    as if the tensor started at word `offset` of a longer array (so per-block
    sums add up to the sum over the concatenation). Chunked to bound the
    temporaries."""
    import torch
    w32 = t.reshape(-1).view(torch.int32)
    total = 0
    for s in range(0, w32.numel(), chunk):
        w = w32[s:s + chunk].long()
        idx = (torch.arange(w.numel(), device=w.device, dtype=torch.long) + offset + s) * 2 + 1
        total += int((w * idx).sum())
    return total % (1 << 64)


def bench_uniform(hip, args, F):
    """C = A * A for A 4096 x 4096 with 16 uniform distinct columns per row,
    U[-1, 1) (the plain CPU-runnable workload): one product is a few
    microseconds of GPU work, so it is measured as an aggregate of
    `--uniform-pairs` independent products (seeds 0, 1, ...) in one batched
    descriptor whose numeric-only computes replay a captured HIP graph, next
    to one product's one-shot time and its numeric-only reuse through a
    single descriptor."""
    import rvsp
    from oracle import oracle
    from rvsp import gen
    from rvsp.batch import Batch, BatchOutput
    n_pairs = args.uniform_pairs
    mats = [gen.uniform_k(4096, 16, seed=s) for s in range(n_pairs)]
    A = mats[0]
    dmats = [rvsp.DeviceCsr.upload(M) for M in mats]
    dA = dmats[0]
    # One product: one-shot (estimation + first compute) and numeric-only.
    oneshot = []
    for _ in range(args.uniform_runs + 1):
        d = rvsp.Descriptor(algo=F.ALGO_HIP_LEARNED, memory=F.MEMORY_DEVICE)
        hip.sync()
        t0 = time.perf_counter()
        nnz = d.estimate(dA, dA)
        dC = rvsp.DeviceCsr(A.rows, A.cols, nnz)
        cm = dC.struct(nnz)
        F.check(d.compute(dA, dA, cm), "compute")
        oneshot.append(time.perf_counter() - t0)
        if _ == 0:
            # Numeric-only reuse on this descriptor (asynchronous launches).
            F.check(d.compute(dA, dA, cm), "compute")
            d.set_async(True)
            reps = 200
            hip.sync()
            t0 = time.perf_counter()
            for _r in range(reps):
                F.check(d.compute(dA, dA, cm), "compute")
            hip.sync()
            t_num_single = (time.perf_counter() - t0) / reps
            got = dC.download()
        d.close()
        dC.free()
    t_oneshot = statistics.median(oneshot[1:])
    rp, ci, v = oracle.spgemm(A, A)
    ok = (got.row_ptr.tobytes() == rp.tobytes() and got.col_idx.tobytes() == ci.tobytes()
          and got.values.tobytes() == v.tobytes())
    err = oracle.f64_error(A, A, rp, ci, got.values)
    products1 = int(np.diff(A.row_ptr)[A.col_idx].sum())
    # Aggregate: every product of the batch, numeric-only, replayed graph.
    bt = Batch(algo=F.ALGO_HIP_LEARNED)
    nnz_b = bt.estimate(dmats, dmats)
    out = BatchOutput([M.rows for M in mats], [M.cols for M in mats], nnz_b)
    for _ in range(3):  # first compute, position maps, capture
        F.check(bt.compute(dmats, dmats, out.structs), "batch compute")
    bt.set_async(True)
    reps = args.uniform_graph_reps
    s = bt.stream()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, s)
    for _ in range(reps):
        F.check(bt.compute(dmats, dmats, out.structs), "batch compute")
    hip.record(e1, s)
    hip.sync()
    t_graph = (time.perf_counter() - t0) / reps
    t_graph_ev = hip.elapsed_ms(e0, e1) * 1e-3 / reps
    graphs = bt.graph_launches()
    upath, u_stream, u_bin = bt.numeric_path()
    last = out.download(n_pairs - 1)
    rp2, ci2, v2 = oracle.spgemm(mats[-1], mats[-1])
    ok &= (last.row_ptr.tobytes() == rp2.tobytes() and last.col_idx.tobytes() == ci2.tobytes()
           and last.values.tobytes() == v2.tobytes())
    out.free()
    bt.close()
    for M in dmats:
        M.free()
    products = sum(int(np.diff(M.row_ptr)[M.col_idx].sum()) for M in mats)
    nnz_total = int(sum(nnz_b))
    alg_bytes = sum(4 * (M.rows + 1) + 8 * M.nnz + 4 * (M.rows + 1) for M in mats) + 8 * nnz_total
    ach = alg_bytes / t_graph_ev / 1e9
    cpu = None
    if not args.no_cpu and oracle.ref_available():
        t_ref = []
        for _ in range(5):
            t0 = time.perf_counter()
            oracle.ref_spgemm(A, A)
            t_ref.append(time.perf_counter() - t0)
        from rvsp import _ffi as FF
        times, _ = oracle.time_reference_numeric(A, A, 20, FF.CsrMatrix)
        cpu = {"value": round(2 * products1 / statistics.median(times) / 1e9, 4),
               "unit": "GFLOP/s", "cores": 1, "kind": "reference",
               "sample": "numeric-only multiply of the seed-0 product through the reference "
                         "descriptor (fill once, untimed), median of 20, one thread",
               "oneshot_ms": round(statistics.median(t_ref) * 1e3, 3),
               "numeric_ms": round(statistics.median(times) * 1e3, 4)}
    return {
        "workload": f"uniform16_n4096_AxA_x{n_pairs}_graph",
        "pairs": n_pairs, "products_per_pair": products1, "products": products,
        "nnz_c": nnz_total,
        "single_product": {"oneshot_ms": round(t_oneshot * 1e3, 3),
                           "numeric_ms": round(t_num_single * 1e3, 4),
                           "numeric_gflops": round(2 * products1 / t_num_single / 1e9, 2)},
        "graph_numeric_ms": round(t_graph * 1e3, 4),
        "graph_numeric_gflops": round(2 * products / t_graph / 1e9, 2),
        "graph_launches": graphs,
        "numeric_path": {"chosen": {1: "persistent", 2: "per_bin"}.get(upath, "none"),
                         "measured_ms": {"persistent": round(u_stream, 3), "per_bin": round(u_bin, 3)}},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "algorithmic_bytes": alg_bytes, "kernel_ms": round(t_graph_ev * 1e3, 4),
                     "scope": "one graph replay (numeric-only multiply of every pair), "
                              "HIP events on the batch's stream"},
        "fp64_error": oracle.merge_f64_errors([err]),
        "verified_bit_exact": bool(ok),
        "cpu_baseline": cpu,
    }


def bench_batched(hip, args, F):
    """2048 independent pairs of uneven shapes: the plain path (one
    descriptor per pair, FIXED) against one batched descriptor (one set of
    launches per phase; numeric-only repeats replayed from a HIP graph).
    Both sides get the same untimed warm-up before their timed numeric-only
    repeats (the first numeric-only compute, which builds the position maps,
    and one more); every run writes outputs pre-filled with a sentinel."""
    import torch
    import rvsp
    from oracle import oracle
    from rvsp import devgen
    from rvsp.batch import Batch
    L = F.lib()
    t0 = time.perf_counter()
    pairs, keep = devgen.lognormal_batch(count=args.pairs, seed=0, device="cuda")
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    products = 0
    for A, B in pairs:
        products += int((B.rp[1:] - B.rp[:-1]).long()[A.ci.long()].sum())
    print(f"# batched: {len(pairs)} pairs, rows {sum(a.rows for a, _ in pairs)}, "
          f"nnz(A)+nnz(B) {sum(a.nnz + b.nnz for a, b in pairs)}, products {products}, "
          f"generated in {t_gen:.1f} s", file=sys.stderr, flush=True)
    n = len(pairs)

    def alloc_out(nnz):
        tot = int(sum(nnz))
        ci = torch.empty(max(tot, 1), dtype=torch.int32, device="cuda")
        val = torch.empty(max(tot, 1), dtype=torch.float32, device="cuda")
        rps = [torch.empty(a.rows + 1, dtype=torch.int32, device="cuda") for a, _ in pairs]
        structs = (F.CsrMatrix * n)()
        off = 0
        for p, (A, B) in enumerate(pairs):
            structs[p] = F.CsrMatrix(A.rows, B.cols, int(nnz[p]), rps[p].data_ptr(),
                                     ci.data_ptr() + 4 * off, val.data_ptr() + 4 * off,
                                     F.DTYPE_FP32, F.FORMAT_CSR, 0)
            off += int(nnz[p])
        return ci, val, rps, structs

    def sentinel(ci, val, rps):
        ci.fill_(-1)
        val.view(torch.int32).fill_(-1)
        for r in rps:
            r.fill_(-1)
        torch.cuda.synchronize()

    def sums(ci, val, rps):
        return (word_checksum(ci), word_checksum(val),
                sum(word_checksum(r) for r in rps) % (1 << 64))

    a_st = [A.struct() for A, _ in pairs]
    b_st = [B.struct() for _, B in pairs]
    reps = args.batch_repeats
    # Plain: a descriptor per pair, each call synchronous.
    descs = []
    hip.sync()
    t0 = time.perf_counter()
    nnz_plain = []
    for p in range(n):
        d = rvsp.Descriptor(algo=F.ALGO_HIP_FIXED, memory=F.MEMORY_DEVICE)
        ws, nz = C.c_size_t(), C.c_int32()
        F.check(L.rvsp_spgemm_work_estimation(d.handle, C.byref(a_st[p]), C.byref(b_st[p]),
                                              C.byref(ws), C.byref(nz)), "estimate")
        nnz_plain.append(nz.value)
        descs.append(d)
    t_plain_est = time.perf_counter() - t0
    ci_o, val_o, rps_o, out = alloc_out(nnz_plain)
    tok = C.c_void_p(descs[0].ws.ctypes.data)

    def plain_all():
        for p in range(n):
            F.check(L.rvsp_spgemm_compute(descs[p].handle, C.byref(a_st[p]), C.byref(b_st[p]),
                                          C.byref(out[p]), tok), "compute")

    sentinel(ci_o, val_o, rps_o)
    hip.sync()
    t0 = time.perf_counter()
    plain_all()
    t_plain_first = time.perf_counter() - t0
    sums_first = sums(ci_o, val_o, rps_o)
    plain_all()  # untimed: first numeric-only compute (position maps built)
    plain_all()  # untimed: second, as the batched side gets
    # Numeric-only computes rewrite the values only (the structure stays from
    # the first compute, checked above): they must rewrite all of them.
    val_o.view(torch.int32).fill_(-1)
    torch.cuda.synchronize()
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_all()
    t_plain_num = (time.perf_counter() - t0) / reps
    sums_plain = sums(ci_o, val_o, rps_o)
    for d in descs:
        d.close()
    print(f"# batched: plain estimate {t_plain_est * 1e3:.1f} ms, first compute "
          f"{t_plain_first * 1e3:.1f} ms, numeric {t_plain_num * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    # Batched.
    bt = Batch(algo=F.ALGO_HIP_LEARNED if args.batch_algo == "learned" else F.ALGO_HIP_FIXED)
    # The operand structs are built once, as the plain side's are.
    As = Batch._structs([A for A, _ in pairs])
    Bs = Batch._structs([B for _, B in pairs])
    hip.sync()
    t0 = time.perf_counter()
    nnz_b = bt.estimate(As, Bs)
    t_b_est = time.perf_counter() - t0
    sentinel(ci_o, val_o, rps_o)
    hip.sync()
    t0 = time.perf_counter()
    F.check(bt.compute(As, Bs, out), "batch compute")
    t_b_first = time.perf_counter() - t0
    same_first = list(nnz_b) == nnz_plain and sums(ci_o, val_o, rps_o) == sums_first
    # Untimed: the first numeric-only compute (maps built; the persistent pass
    # timed against the per-bin launches, the faster kept), then the capture.
    F.check(bt.compute(As, Bs, out), "batch compute")
    F.check(bt.compute(As, Bs, out), "batch compute")
    path, t_path_stream, t_path_bin = bt.numeric_path()
    val_o.view(torch.int32).fill_(-1)  # the replays rewrite every value
    torch.cuda.synchronize()
    bt.set_async(True)
    s_b = bt.stream()
    e0, e1 = hip.event(), hip.event()
    hip.sync()
    t0 = time.perf_counter()
    hip.record(e0, s_b)
    for _ in range(reps):
        F.check(bt.compute(As, Bs, out), "batch compute")
    hip.record(e1, s_b)
    hip.sync()
    t_b_wall = (time.perf_counter() - t0) / reps
    t_b_num = hip.elapsed_ms(e0, e1) * 1e-3 / reps
    groups, graphs = bt.groups(), bt.graph_launches()
    bt.close()
    same = same_first and sums(ci_o, val_o, rps_o) == sums_plain
    # Oracle (bits) and FP64 error on a few pairs.
    oracle_ok = True
    errs = []
    for p in range(0, n, max(1, n // 4)):
        A, B = pairs[p][0].to_host(), pairs[p][1].to_host()
        rp, ci, v = oracle.spgemm(A, B)
        o = out[p]
        k = int(nnz_b[p])
        base = (o.col_idx - ci_o.data_ptr()) // 4
        got_rp = rps_o[p].cpu().numpy()
        got_ci = ci_o[base:base + k].cpu().numpy()
        got_v = val_o[base:base + k].cpu().numpy()
        oracle_ok &= (got_rp.tobytes() == rp.tobytes() and got_ci.tobytes() == ci.tobytes()
                      and got_v.tobytes() == v.tobytes())
        errs.append(oracle.f64_error(A, B, rp, ci, got_v))
    cpu = None
    if not args.no_cpu:
        host_pairs = ((A.to_host(), B.to_host()) for A, B in pairs)
        spent, prods, npairs, kind = cpu_oneshot(host_pairs, args.cpu_seconds / 2)
        cpu = {"value": round(2 * prods / spent / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
               "kind": kind, "sample": f"one-shot products of the first {npairs} pairs"}
    nnz_total = int(sum(nnz_b))
    del ci_o, val_o, rps_o, keep
    torch.cuda.empty_cache()
    flops = 2.0 * products
    oneshot_plain = t_plain_est + t_plain_first
    oneshot_b = t_b_est + t_b_first
    # Algorithmic bytes of one numeric-only multiply of every pair (A and B
    # read, C's structure read, C's values written): A != B in every pair.
    alg_bytes = sum(4 * (A.rows + 1) + 8 * A.nnz + 4 * (B.rows + 1) + 8 * B.nnz
                    + 4 * (A.rows + 1) for A, B in pairs) + 8 * nnz_total
    ach = alg_bytes / t_b_num / 1e9
    return {
        "workload": f"lognormal_pairs{n}_n256-65536",
        "pairs": n, "products": products, "nnz_c": nnz_total, "groups": groups,
        "algo_batched": args.batch_algo,
        "oneshot_ms": {"plain": round(oneshot_plain * 1e3, 2), "batched": round(oneshot_b * 1e3, 2)},
        "oneshot_phases_ms": {"plain": {"estimate": round(t_plain_est * 1e3, 2),
                                        "first_compute": round(t_plain_first * 1e3, 2)},
                              "batched": {"estimate": round(t_b_est * 1e3, 2),
                                          "first_compute": round(t_b_first * 1e3, 2)}},
        "oneshot_gflops": {"plain": round(flops / oneshot_plain / 1e9, 2),
                           "batched": round(flops / oneshot_b / 1e9, 2)},
        "speedup_oneshot": round(oneshot_plain / oneshot_b, 3),
        "numeric_ms": {"plain": round(t_plain_num * 1e3, 3), "batched_graph": round(t_b_num * 1e3, 3),
                       "batched_graph_wall": round(t_b_wall * 1e3, 3)},
        "numeric_path": {"chosen": {1: "persistent", 2: "per_bin"}.get(path, "none"),
                         "measured_ms": {"persistent": round(t_path_stream, 3),
                                         "per_bin": round(t_path_bin, 3)}},
        "plain_ms_per_call": round(t_plain_num * 1e3 / n, 4),
        "numeric_gflops": {"plain": round(flops / t_plain_num / 1e9, 2),
                           "batched_graph": round(flops / t_b_num / 1e9, 2)},
        "speedup_numeric": round(t_plain_num / t_b_num, 3),
        "graph_launches": graphs,
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "algorithmic_bytes": alg_bytes,
                     "scope": "numeric-only multiply of all pairs (graph replay, HIP events on "
                              "the batch's stream)"},
        "fp64_error": oracle.merge_f64_errors(errs),
        "verified_bit_exact": bool(same and oracle_ok),
        "cpu_baseline": cpu,
    }


def row_blocks(A, max_products):
    """Row cuts of A so that every block of C = A*A has at most
    max_products products."""
    lens = np.diff(A.row_ptr).astype(np.int64)
    cs = np.concatenate([[0], np.cumsum(lens[A.col_idx])])
    ops = cs[A.row_ptr[1:]] - cs[A.row_ptr[:-1]]
    cuts, acc = [0], 0
    for i, o in enumerate(ops):
        if acc + o > max_products and i > cuts[-1]:
            cuts.append(i)
            acc = 0
        acc += int(o)
    cuts.append(A.rows)
    return cuts, ops


def spread(xs):
    """(median, relative spread (max - min) / median) of a list of times."""
    m = statistics.median(xs)
    return m, (max(xs) - min(xs)) / m if m > 0 else 0.0


def bench_powerlaw(hip, args, F):
    """R-MAT C = A*A: nnz(C) exceeds int32, so the product runs as row
    blocks of A against the same B. Plain: one FIXED descriptor per block,
    one after another. Batched: all blocks as pairs of one batched
    descriptor, with the FIXED table and with the LEARNED selector. Each
    plan runs once untimed, then `--powerlaw-runs` timed one-shot runs
    (estimation + compute, fresh descriptors, interleaved between plans);
    every run writes outputs pre-filled with a sentinel, and each plan's
    last run is checksummed against the others."""
    import torch
    import rvsp
    from oracle import oracle
    from rvsp import gen
    from rvsp.batch import Batch
    L = F.lib()
    t0 = time.perf_counter()
    A = gen.rmat(args.rmat_scale, seed=0)
    cuts, ops = row_blocks(A, 1 << 30)
    products = int(ops.sum())
    print(f"# powerlaw: R-MAT 2^{args.rmat_scale} nnz {A.nnz}, products {products}, "
          f"{len(cuts) - 1} row blocks, built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    dev = "cuda"
    rp = torch.as_tensor(A.row_ptr.astype(np.int32), device=dev)
    ci = torch.as_tensor(A.col_idx.astype(np.int32), device=dev)
    val = torch.as_tensor(A.values, device=dev)
    full = F.CsrMatrix(A.rows, A.cols, A.nnz, rp.data_ptr(), ci.data_ptr(), val.data_ptr(),
                       F.DTYPE_FP32, F.FORMAT_CSR, 0)
    blocks, brps = [], []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        brp = (rp[r0:r1 + 1] - rp[r0]).contiguous()
        brps.append(brp)
        o = int(A.row_ptr[r0])
        blocks.append(F.CsrMatrix(r1 - r0, A.cols, int(A.row_ptr[r1]) - o, brp.data_ptr(),
                                  ci.data_ptr() + 4 * o, val.data_ptr() + 4 * o,
                                  F.DTYPE_FP32, F.FORMAT_CSR, 0))
    nb = len(blocks)
    # nnz of every block (untimed), then one output laid out block after block.
    nnz_blocks = []
    for p in range(nb):
        d = rvsp.Descriptor(algo=F.ALGO_HIP_FIXED, memory=F.MEMORY_DEVICE)
        ws, nz = C.c_size_t(), C.c_int32()
        F.check(L.rvsp_spgemm_work_estimation(d.handle, C.byref(blocks[p]), C.byref(full),
                                              C.byref(ws), C.byref(nz)), "estimate")
        nnz_blocks.append(nz.value)
        d.close()
    total = int(sum(nnz_blocks))
    torch.cuda.empty_cache()
    c_rps = [torch.empty(b.rows + 1, dtype=torch.int32, device=dev) for b in blocks]
    ci_out = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    val_out = torch.empty(max(total, 1), dtype=torch.float32, device=dev)
    structs = (F.CsrMatrix * nb)()
    off = 0
    for p in range(nb):
        structs[p] = F.CsrMatrix(blocks[p].rows, A.cols, nnz_blocks[p], c_rps[p].data_ptr(),
                                 ci_out.data_ptr() + 4 * off, val_out.data_ptr() + 4 * off,
                                 F.DTYPE_FP32, F.FORMAT_CSR, 0)
        off += nnz_blocks[p]

    def sentinel():
        ci_out.fill_(-1)
        val_out.view(torch.int32).fill_(-1)
        for r in c_rps:
            r.fill_(-1)
        torch.cuda.synchronize()

    def run_plain():
        est = comp = 0.0
        for p in range(nb):
            d = rvsp.Descriptor(algo=F.ALGO_HIP_FIXED, memory=F.MEMORY_DEVICE)
            ws, nz = C.c_size_t(), C.c_int32()
            hip.sync()
            t0 = time.perf_counter()
            F.check(L.rvsp_spgemm_work_estimation(d.handle, C.byref(blocks[p]), C.byref(full),
                                                  C.byref(ws), C.byref(nz)), "estimate")
            t1 = time.perf_counter()
            if nz.value != nnz_blocks[p]:
                raise RuntimeError(f"block {p}: nnz {nz.value} != {nnz_blocks[p]}")
            cm = structs[p]
            F.check(L.rvsp_spgemm_compute(d.handle, C.byref(blocks[p]), C.byref(full),
                                          C.byref(cm), C.c_void_p(d.ws.ctypes.data)), "compute")
            t2 = time.perf_counter()
            est += t1 - t0
            comp += t2 - t1
            d.close()
        return est, comp, nb

    def run_batched(algo):
        bt = Batch(algo=algo)
        hip.sync()
        t0 = time.perf_counter()
        nnz_b = bt.estimate(list(blocks), [full] * nb)
        t1 = time.perf_counter()
        F.check(bt.compute(list(blocks), [full] * nb, structs), "batch compute")
        t2 = time.perf_counter()
        groups = bt.groups()
        bt.close()
        if list(nnz_b) != nnz_blocks:
            raise RuntimeError("batched nnz differs from the plain path")
        return t1 - t0, t2 - t1, groups

    def sums():
        out = (word_checksum(ci_out[:total]), word_checksum(val_out[:total]),
               sum(word_checksum(r) for r in c_rps) % (1 << 64))
        # The checksums' temporaries go back to the device, so the next
        # plan's allocations find the memory the earlier runs left them.
        torch.cuda.empty_cache()
        return out

    plans = {"plain_fixed": run_plain,
             "batched_fixed": lambda: run_batched(F.ALGO_HIP_FIXED),
             "learned_batched": lambda: run_batched(F.ALGO_HIP_LEARNED)}
    if args.powerlaw_plans:  # profiling runs: a subset of the plans
        plans = {k: v for k, v in plans.items() if k in args.powerlaw_plans.split(",")}
    for name, fn in plans.items():  # untimed: first use of every kernel, model, stream
        sentinel()
        fn()
    runs = {k: [] for k in plans}
    checks = {}
    for r in range(args.powerlaw_runs):
        for name, fn in plans.items():
            sentinel()
            est, comp, groups = fn()
            runs[name].append((est, comp, groups))
            print(f"# powerlaw run {r}: {name} {(est + comp) * 1e3:.1f} ms (estimate "
                  f"{est * 1e3:.1f}, compute {comp * 1e3:.1f}, {groups} groups)",
                  file=sys.stderr, flush=True)
            if r == args.powerlaw_runs - 1:
                checks[name] = sums()
    same = len(set(checks.values())) == 1
    # Oracle (FP32, bit for bit) and FP64 error on four random ranges of 64
    # rows of the last run's output.
    oracle_ok = True
    errs = []
    rng = np.random.default_rng(1)
    starts = rng.integers(0, A.rows - 64, size=4)
    off_blk = np.concatenate([[0], np.cumsum(nnz_blocks)])
    for r0 in starts:
        b = int(np.searchsorted(cuts, r0, side="right") - 1)
        r1 = min(int(r0) + 64, cuts[b + 1])
        want_rp, want_ci, want_v = oracle.spgemm(A, A, int(r0), r1)
        crp = c_rps[b].cpu().numpy()
        lo, hi = int(r0) - cuts[b], r1 - cuts[b]
        s, e = int(crp[lo]), int(crp[hi])
        got_ci = ci_out[off_blk[b] + s:off_blk[b] + e].cpu().numpy()
        got_v = val_out[off_blk[b] + s:off_blk[b] + e].cpu().numpy()
        oracle_ok &= (np.array_equal(crp[lo:hi + 1] - s, want_rp) and
                      got_ci.tobytes() == want_ci.tobytes() and got_v.tobytes() == want_v.tobytes())
        errs.append(oracle.f64_error(A, A, want_rp, want_ci, got_v, int(r0)))
    cpu = None
    if not args.no_cpu:
        # Reference one-shot on whole row blocks of 256 rows spread over A.
        sample_rows = np.linspace(0, A.rows - 256, 24).astype(np.int64)
        spent, prods = 0.0, 0
        from rvsp.api import Csr
        kind = "reference" if oracle.ref_available() else "port"
        for r0 in sample_rows:
            r1 = int(r0) + 256
            o0, o1 = A.row_ptr[r0], A.row_ptr[r1]
            blk = Csr(256, A.cols, A.row_ptr[r0:r1 + 1] - o0, A.col_idx[o0:o1], A.values[o0:o1])
            t0 = time.perf_counter()
            if kind == "reference":
                oracle.ref_spgemm(blk, A)
            else:
                oracle.spgemm(blk, A)
            spent += time.perf_counter() - t0
            prods += int(ops[r0:r1].sum())
            if spent > args.cpu_seconds / 2:
                break
        cpu = {"value": round(2 * prods / spent / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
               "kind": kind, "sample": "one-shot products of 256-row blocks spread over A"}
    del ci_out, val_out, c_rps
    torch.cuda.empty_cache()
    flops = 2.0 * products
    med = {k: spread([e + c for e, c, _ in v]) for k, v in runs.items()}
    ms = {k: round(m * 1e3, 1) for k, (m, _) in med.items()}
    ratio = (lambda a, b: round(med[a][0] / med[b][0], 3) if a in med and b in med else None)
    lb = med["learned_batched"][0] if "learned_batched" in med else min(m for m, _ in med.values())
    # Algorithmic bytes of the whole product (A = B counted once; C's row
    # pointers once per row plus one per block).
    alg_bytes = 4 * (A.rows + 1) + 8 * A.nnz + 4 * (A.rows + nb) + 8 * total
    ach = alg_bytes / lb / 1e9
    return {
        "workload": f"rmat_s{args.rmat_scale}_ef16_AxA_rowblocks",
        "products": products, "nnz_c": total, "row_blocks": nb,
        "runs": args.powerlaw_runs,
        "oneshot_ms": ms,
        "spread": {k: round(sp, 3) for k, (_, sp) in med.items()},
        "phases_ms": {k: {"estimate": round(statistics.median(e for e, _, _ in v) * 1e3, 1),
                          "compute": round(statistics.median(c for _, c, _ in v) * 1e3, 1)}
                      for k, v in runs.items()},
        "all_runs_ms": {k: [round((e + c) * 1e3, 1) for e, c, _ in v] for k, v in runs.items()},
        "groups": {k: v[-1][2] for k, v in runs.items()},
        "gflops": {k: round(flops / m / 1e9, 2) for k, (m, _) in med.items()},
        "speedup_learned_batched_vs_plain": ratio("plain_fixed", "learned_batched"),
        "speedup_learned_vs_fixed_batched": ratio("batched_fixed", "learned_batched"),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "algorithmic_bytes": alg_bytes, "scope": "whole one-shot product "
                     "(learned_batched, estimation + compute)"},
        "fp64_error": oracle.merge_f64_errors(errs),
        "verified_bit_exact": bool(same and oracle_ok),
        "cpu_baseline": cpu,
    }


def bench_banded(hip, args, F):
    """A 500k x 200k banded (24 of 32) with every 100th row dense-ish (2000),
    B 200k x 500k uniform 8 per row: FIXED vs LEARNED, one-shot and
    numeric-only."""
    import rvsp
    from oracle import oracle
    from rvsp import gen
    L = F.lib()
    A, B = gen.banded_mixed(seed=0)
    products = int(np.diff(B.row_ptr)[A.col_idx].sum())
    print(f"# banded: A {A.rows}x{A.cols} nnz {A.nnz}, B nnz {B.nnz}, products {products}",
          file=sys.stderr, flush=True)
    import torch
    from rvsp.api import Csr
    dA, dB = rvsp.DeviceCsr.upload(A), rvsp.DeviceCsr.upload(B)
    res = {}
    outs = {}
    c_rp = torch.empty(A.rows + 1, dtype=torch.int32, device="cuda")
    c_ci = c_val = None
    for name, algo in (("fixed", F.ALGO_HIP_FIXED), ("learned", F.ALGO_HIP_LEARNED)):
        oneshot, phases = [], []
        for run in range(args.banded_runs + 1):  # run 0 untimed: first use
            d = rvsp.Descriptor(algo=algo, memory=F.MEMORY_DEVICE)
            hip.sync()
            t0 = time.perf_counter()
            nnz = d.estimate(dA, dB)
            t1 = time.perf_counter()
            if c_ci is None:
                c_ci = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
                c_val = torch.empty(max(nnz, 1), dtype=torch.float32, device="cuda")
            for t in (c_rp, c_ci, c_val.view(torch.int32)):
                t.fill_(-1)
            torch.cuda.synchronize()
            cm = F.CsrMatrix(A.rows, B.cols, nnz, c_rp.data_ptr(), c_ci.data_ptr(),
                             c_val.data_ptr(), F.DTYPE_FP32, F.FORMAT_CSR, 0)
            hip.sync()
            t2 = time.perf_counter()
            F.check(d.compute(dA, dB, cm), "compute")
            t3 = time.perf_counter()
            if run:
                oneshot.append((t1 - t0) + (t3 - t2))
                phases.append((t1 - t0, t3 - t2))
            if run < args.banded_runs:
                d.close()
        # Numeric-only repeats on the last descriptor: one untimed (position
        # maps), then values rewritten from the sentinel by the timed ones.
        F.check(d.compute(dA, dB, cm), "compute")
        c_val.view(torch.int32).fill_(-1)
        torch.cuda.synchronize()
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(20):
            F.check(d.compute(dA, dB, cm), "compute")
        num = (time.perf_counter() - t0) / 20
        d.close()
        res[name] = {"oneshot_ms": statistics.median(oneshot) * 1e3, "numeric_ms": num * 1e3,
                     "oneshot_spread": spread(oneshot)[1],
                     "estimate_ms": statistics.median(e for e, _ in phases) * 1e3,
                     "first_compute_ms": statistics.median(c for _, c in phases) * 1e3}
        outs[name] = Csr(A.rows, B.cols, c_rp.cpu().numpy(), c_ci[:nnz].cpu().numpy(),
                         c_val[:nnz].cpu().numpy())
    same = all(getattr(outs["fixed"], k).tobytes() == getattr(outs["learned"], k).tobytes()
               for k in ("row_ptr", "col_idx", "values"))
    rp, ci, v = oracle.spgemm(A, B, 0, 5000)
    got = outs["fixed"]
    ok = (np.array_equal(got.row_ptr[:5001], rp) and got.col_idx[:rp[-1]].tobytes() == ci.tobytes()
          and got.values[:rp[-1]].tobytes() == v.tobytes())
    err = oracle.f64_error(A, B, rp, ci, got.values[:rp[-1]])
    dA.free()
    dB.free()
    del c_rp, c_ci, c_val
    flops = 2.0 * products
    nnz_c = int(outs["fixed"].nnz)
    alg_bytes = (4 * (A.rows + 1) + 8 * A.nnz + 4 * (B.rows + 1) + 8 * B.nnz
                 + 4 * (A.rows + 1) + 8 * nnz_c)
    best = min(r["numeric_ms"] for r in res.values()) * 1e-3
    ach = alg_bytes / best / 1e9
    return {
        "workload": "banded500k_x_uniform200k_mixed_dense_rows",
        "products": products, "nnz_c": int(outs["fixed"].nnz),
        "oneshot_ms": {k: round(r["oneshot_ms"], 2) for k, r in res.items()},
        "oneshot_phases_ms": {k: {"estimate": round(r["estimate_ms"], 2),
                                  "first_compute": round(r["first_compute_ms"], 2)}
                              for k, r in res.items()},
        "oneshot_spread": {k: round(r["oneshot_spread"], 3) for k, r in res.items()},
        "runs": args.banded_runs,
        "numeric_ms": {k: round(r["numeric_ms"], 3) for k, r in res.items()},
        "numeric_gflops": {k: round(flops / (r["numeric_ms"] * 1e-3) / 1e9, 2) for k, r in res.items()},
        "speedup_learned_vs_fixed_oneshot": round(res["fixed"]["oneshot_ms"] / res["learned"]["oneshot_ms"], 3),
        "speedup_learned_vs_fixed_numeric": round(res["fixed"]["numeric_ms"] / res["learned"]["numeric_ms"], 3),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "algorithmic_bytes": alg_bytes,
                     "scope": "numeric-only multiply (faster plan)"},
        "fp64_error": oracle.merge_f64_errors([err]),
        "verified_bit_exact": bool(same and ok),
    }


def bench_legacy(hip, args, F):
    """FP64 -> FP64 and INT8 -> INT32 one-shot products (the reference's
    legacy kernels) of the 5-point Laplacian with itself through
    rvsp_spgemm_csr on host arrays: the GPU call end to end (copies in and
    out included) against the reference's own scalar legacy kernel, or the
    oracle's restatement, on one host core."""
    import rvsp
    from oracle import oracle
    from rvsp import gen
    from rvsp.api import Csr
    A = gen.laplacian_2d(args.grid)
    products = int(np.diff(A.row_ptr)[A.col_idx].sum())
    rng = np.random.default_rng(0)
    out = {"workload": f"laplacian5pt_g{args.grid}_AxA_oneshot_legacy_dtypes", "products": products}
    for name, vals, dt in (
            ("fp64", rng.uniform(0.5, 1.5, size=A.nnz), np.float64),
            ("int8", rng.integers(-128, 128, size=A.nnz), np.int8)):
        M = Csr(A.rows, A.cols, A.row_ptr, A.col_idx, vals, dtype=dt)
        rvsp.spgemm_legacy(M, M)  # warm-up
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            got = rvsp.spgemm_legacy(M, M)
            times.append(time.perf_counter() - t0)
        t_gpu = statistics.median(times)
        kind = "reference" if oracle.ref_available() else "port"
        t0 = time.perf_counter()
        if kind == "reference":
            st, want = oracle.ref_spgemm_legacy(M, M)
        else:
            want = oracle.spgemm_legacy(M, M)
        t_cpu = time.perf_counter() - t0
        same = all(g.tobytes() == w.tobytes()
                   for g, w in zip((got.row_ptr, got.col_idx, got.values), want))
        out[name] = {"oneshot_ms": round(t_gpu * 1e3, 2), "nnz_c": int(got.nnz),
                     "gflops": round(2.0 * products / t_gpu / 1e9, 2),
                     "cpu_ms": round(t_cpu * 1e3, 1), "cpu_kind": kind, "cpu_cores": 1,
                     "speedup_vs_cpu": round(t_cpu / t_gpu, 2), "verified_bit_exact": bool(same)}
        print(f"# legacy {name}: GPU one-shot {t_gpu * 1e3:.2f} ms, CPU ({kind}) "
              f"{t_cpu * 1e3:.1f} ms, bit-exact {same}", file=sys.stderr, flush=True)
    return out


def bench_headline(hip, args, F):
    """The headline: numeric-only multiplies of the Laplacian product."""
    import rvsp
    from rvsp import gen
    A = gen.laplacian_2d(args.grid)
    n_vals = args.steps + args.warmup
    print(f"# laplacian5pt g={args.grid}: rows={A.rows} nnz={A.nnz}; staging {n_vals} "
          f"value sets", file=sys.stderr, flush=True)
    host_vals = [gen.laplacian_values(A, i) for i in range(n_vals)]
    dA = rvsp.DeviceCsr.upload(A.with_values(host_vals[0]))
    val_bufs = []
    for v in host_vals:
        p = rvsp.device_alloc(v.nbytes)
        rvsp.to_device(p, v)
        val_bufs.append(p)

    runs = {}
    if args.only in (None, "fixed"):
        runs["fixed"] = run_plan(hip, A, dA, val_bufs, F.ALGO_HIP_FIXED, args.steps,
                                 args.warmup, 0)
        print(f"# fixed: {runs['fixed']['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    if args.only in (None, "tuned"):
        runs["tuned"] = run_plan(hip, A, dA, val_bufs, F.ALGO_HIP_TUNED, args.steps,
                                 args.warmup, args.autotune_reps)
        print(f"# tuned: {runs['tuned']['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    head = runs.get("tuned") or runs["fixed"]
    head_name = "tuned" if "tuned" in runs else "fixed"

    verified = None
    fp64 = None
    if not args.no_verify:
        from oracle import oracle
        last = A.with_values(host_vals[args.steps - 1])
        rp, ci, val = oracle.spgemm(last, last)
        verified = True
        for r in runs.values():
            got = r["C"].download()
            verified &= (got.row_ptr.tobytes() == rp.tobytes()
                         and got.col_idx.tobytes() == ci.tobytes()
                         and got.values.tobytes() == val.tobytes())
        fp64 = oracle.merge_f64_errors([oracle.f64_error(last, last, rp, ci, got.values)])

    ops = rvsp.Descriptor(memory=F.MEMORY_DEVICE)
    ops.estimate(dA, dA)
    products = int(ops.op_counts(A.rows).sum())
    ops.close()
    flops = 2.0 * products
    M, nnz_c = A.rows, head["nnz_c"]
    alg_bytes = 4 * (M + 1) + 8 * A.nnz + 4 * (M + 1) + 8 * nnz_c
    achieved = alg_bytes / (head["kernel_ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        prof = json.load(open(PROFILE_TRAFFIC)).get(f"laplacian5pt_g{args.grid}")
        if prof:
            traffic = prof.get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        med_s, reps, kind = cpu_baseline(A, host_vals[:max(3, min(len(host_vals), 64))],
                                         args.cpu_seconds)
        cpu = {"value": round(flops / med_s / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
               "kind": kind,
               "sample": f"{reps} numeric-only multiplies of this workload through the "
                         f"descriptor (fill once, untimed), median, one thread",
               "ms_per_multiply": round(med_s * 1e3, 3)}
        from oracle import oracle
        if kind == "reference":
            # All cores: row blocks balanced by products, one reference
            # descriptor and one host thread each, for a few seconds.
            threads = max(1, min(16, os.cpu_count() or 1))
            try:
                rate, used = oracle.time_reference_numeric_threads(
                    A, A, threads, min(5.0, args.cpu_seconds / 2), F.CsrMatrix)
                cpu["all_cores"] = {"value": round(2 * rate / 1e9, 4), "unit": "GFLOP/s",
                                    "cores": used,
                                    "sample": "numeric-only multiplies of row blocks, one "
                                              "descriptor and one thread per block, 5 s"}
            except Exception as e:  # reported, never hidden
                cpu["all_cores"] = {"error": f"{type(e).__name__}: {e}"}

    def gflops(r):
        return round(flops / (r["ms_per_step"] * 1e-3) / 1e9, 3)

    out = {
        "metric": "SpGEMM GFLOP/s (C=A*A numeric-only multiply, descriptor reuse, 1 GPU)",
        "value": gflops(head),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (5-point Laplacian, values U[0.5,1.5) redrawn per multiply)",
        "config": {"workload": f"laplacian5pt_g{args.grid}_AxA_numeric_reuse",
                   "rows": M, "nnz_a": A.nnz, "nnz_c": nnz_c, "products": products,
                   "plan": head_name},
        "hbm_gbps": round(achieved, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic, "kernel_ms": round(head["kernel_ms"], 5),
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
        "plans": {k: {"gflops": gflops(r), "ms_per_step": round(r["ms_per_step"], 5),
                      "kernel_ms": round(r["kernel_ms"], 5),
                      "numeric_bins": [{q: p[q] for q in ("kind", "group", "cap", "rows")}
                                       for p in r["plan"]]}
                  for k, r in runs.items()},
        "speedup_tuned_vs_fixed": (round(runs["fixed"]["ms_per_step"] / runs["tuned"]["ms_per_step"], 3)
                                   if len(runs) == 2 else None),
        "phases_ms": {"work_estimation": round(head["t_sym"] * 1e3, 3),
                      "first_compute_fill_numeric": round(head["t_first"] * 1e3, 3),
                      "autotune": round(head["t_tune"] * 1e3, 1) if head["t_tune"] else None},
        "verified_bit_exact": verified,
        "fp64_error": fp64,
    }
    for r in runs.values():
        r["C"].free()
    for p in val_bufs:
        rvsp.device_free(p)
    dA.free()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--autotune-reps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--only", choices=["fixed", "tuned"], default=None,
                    help="time one plan only (profiling runs)")
    ap.add_argument("--extra", default="uniform,batched,powerlaw,banded",
                    help="secondary workloads reported under 'configs' (comma list from "
                         "uniform, batched, powerlaw, banded, legacy; or none)")
    ap.add_argument("--only-extra", action="store_true",
                    help="skip the headline workload (profiling runs of the extras)")
    ap.add_argument("--pairs", type=int, default=2048)
    ap.add_argument("--uniform-pairs", type=int, default=64)
    ap.add_argument("--uniform-runs", type=int, default=5)
    ap.add_argument("--uniform-graph-reps", type=int, default=50)
    ap.add_argument("--batch-repeats", type=int, default=5)
    ap.add_argument("--batch-algo", choices=["fixed", "learned"], default="learned")
    ap.add_argument("--rmat-scale", type=int, default=20)
    ap.add_argument("--banded-runs", type=int, default=3)
    ap.add_argument("--powerlaw-runs", type=int, default=5)
    ap.add_argument("--powerlaw-plans", default=None,
                    help="comma list of plain_fixed, batched_fixed, learned_batched "
                         "(profiling runs; default all)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (a known "
                         "byte count for scaling the HBM traffic counters), then exit")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        import rvsp
        n = args.calibrate_copy_mib << 20
        src, dst = rvsp.device_alloc(n), rvsp.device_alloc(n)
        for _ in range(3):
            rvsp._ffi.check(rvsp.lib().rvsp_hip_memcpy(C.c_void_p(dst), C.c_void_p(src), n, 3))
        rvsp.synchronize()
        print(json.dumps({"calibration_copy_bytes": n, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the SpGEMM path has no multi-GPU mode. It "
              "would shard by output rows: each GPU runs this pipeline on a block of "
              "A's rows balanced by flop count, with a replica of B.", file=sys.stderr)
        sys.exit(2)

    extra = [x for x in args.extra.split(",") if x and x != "none"]
    if extra:
        # The secondary workloads build operands with torch. torch brings its
        # own HIP runtime; loading it first makes the library bind to that
        # same runtime (one runtime per process).
        import torch
        torch.zeros(1, device="cuda")
    import rvsp
    from rvsp import _ffi as F

    hip = Hip()
    out = {} if args.only_extra else bench_headline(hip, args, F)
    extras = {"uniform": bench_uniform, "batched": bench_batched, "powerlaw": bench_powerlaw, "banded": bench_banded,
              "legacy": bench_legacy}
    configs = {}
    for name in extra:
        # Each workload starts from an empty device buffer cache (its own
        # untimed first runs fill it), so what one workload left cached does
        # not decide whether the next one's timed runs allocate.
        F.check(F.lib().rvsp_hip_pool_release(), "pool_release")
        t0 = time.perf_counter()
        try:
            configs[name] = extras[name](hip, args, F)
        except Exception as e:  # reported, never hidden: the line says which failed
            configs[name] = {"error": f"{type(e).__name__}: {e}"}
        configs[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        print(f"# {name}: {json.dumps(configs[name])}", file=sys.stderr, flush=True)
    out["configs"] = configs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
