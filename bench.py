#!/usr/bin/env python3
"""SpGEMM throughput on one MI355X: GFLOP/s and HBM GB/s.

Workload (default): C = A * A for the 5-point Laplacian of a 1024 x 1024
grid (1,048,576 rows, FP32), structure analysed once by the descriptor, then
numeric-only multiplies with A's values redrawn U[0.5, 1.5) per multiply
(seed = multiply index). All value sets are staged in HBM before the timed
region; a step is one numeric-only multiply through rvsp_spgemm_compute on
device-resident operands. GFLOP/s = 2 * intermediate products / time, the
reference's definition (bench/bench.c:753-754).

Prints one JSON line. `roofline` covers the numeric kernel: algorithmic
bytes per multiply (4(M+1) + 8 nnz(A) + 4(M+1) + 8 nnz(C), A counted once)
over its average duration from HIP events on the descriptor's stream.
`cpu_baseline` times the reference's own descriptor path (built from its
sources into oracle/_ref) on the host, one thread, on a bounded number of
the same multiplies; the project's oracle restatement stands in where the
reference library is absent.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rv-sparse_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

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
    reps = 3
    if oracle.ref_available():
        probe, nnz = oracle.time_reference_numeric(A, A, 1, F.CsrMatrix, values[:1])
        reps = max(3, min(len(values), int(budget_s / max(probe[0], 1e-6))))
        times, _ = oracle.time_reference_numeric(A, A, reps, F.CsrMatrix, values)
        kind = "reference"
    else:
        probe, _ = oracle.time_port_numeric(A, A, 1)
        reps = max(3, int(budget_s / max(probe[0], 1e-6)))
        times, _ = oracle.time_port_numeric(A, A, reps)
        kind = "port"
    return statistics.median(times), reps, kind


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--algo", default="tuned", choices=["fixed", "tuned", "learned"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X, and "
              "the SpGEMM path has no multi-GPU mode (see DESIGN.md for how it would "
              "shard by output rows).", file=sys.stderr)
        sys.exit(2)

    import rvsp
    from rvsp import _ffi as F
    from rvsp import gen

    algo = {"fixed": F.ALGO_HIP_FIXED, "tuned": F.ALGO_HIP_TUNED,
            "learned": F.ALGO_HIP_LEARNED}[args.algo]
    hip = Hip()
    A = gen.laplacian_2d(args.grid)
    n_steps = args.steps + args.warmup
    print(f"# workload laplacian5pt g={args.grid}: rows={A.rows} nnz={A.nnz}; "
          f"staging {n_steps} value sets", file=sys.stderr, flush=True)
    host_vals = [gen.laplacian_values(A, i) for i in range(n_steps)]
    dA = rvsp.DeviceCsr.upload(A.with_values(host_vals[0]))
    val_bufs = []
    for v in host_vals:
        p = rvsp.device_alloc(v.nbytes)
        rvsp.to_device(p, v)
        val_bufs.append(p)

    d = rvsp.Descriptor(algo=algo, memory=F.MEMORY_DEVICE)
    hip.sync()
    t0 = time.perf_counter()
    nnz_c = d.estimate(dA, dA)
    t_sym = time.perf_counter() - t0
    products = int(d.op_counts(A.rows).sum())
    dC = rvsp.DeviceCsr(A.rows, A.cols, nnz_c)
    Cm = dC.struct(nnz_c)
    t0 = time.perf_counter()
    F.check(d.compute(dA, dA, Cm), "first compute")
    t_first = time.perf_counter() - t0
    plan = [p for p in d.plan() if p["phase"] == F.PHASE_NUMERIC and p["rows"] > 0]

    d.set_async(True)
    stream = d.stream()
    a_struct = dA.struct()

    ws = C.c_void_p(d.ws.ctypes.data)

    def step(i):
        # A and B are the same device matrix; only the values buffer changes.
        a_struct.values = val_bufs[i]
        return F.lib().rvsp_spgemm_compute(d.handle, C.byref(a_struct), C.byref(a_struct),
                                           C.byref(Cm), ws)

    for i in range(args.warmup):
        F.check(step(args.steps + i), "warmup")
    ev0 = [hip.event() for _ in range(args.steps)]
    ev1 = [hip.event() for _ in range(args.steps)]
    hip.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        hip.record(ev0[i], stream)
        st = step(i)
        hip.record(ev1[i], stream)
        if st != 0:
            F.check(st, "compute")
    hip.sync()
    wall = time.perf_counter() - t0
    kernel_ms = [hip.elapsed_ms(ev0[i], ev1[i]) for i in range(args.steps)]

    verified = None
    if not args.no_verify:
        from oracle import oracle
        last = A.with_values(host_vals[args.steps - 1])
        rp, ci, val = oracle.spgemm(last, last)
        got = dC.download()
        verified = (got.row_ptr.tobytes() == rp.tobytes() and got.col_idx.tobytes() == ci.tobytes()
                    and got.values.tobytes() == val.tobytes())

    ms_per_step = wall * 1e3 / args.steps
    flops = 2.0 * products
    avg_kernel_ms = statistics.mean(kernel_ms)
    M = A.rows
    alg_bytes = 4 * (M + 1) + 8 * A.nnz + 4 * (M + 1) + 8 * nnz_c
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        prof = json.load(open(PROFILE_TRAFFIC)).get(f"laplacian5pt_g{args.grid}_{args.algo}")
        if prof:
            traffic = prof.get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        med_s, reps, kind = cpu_baseline(A, host_vals[:max(3, min(len(host_vals), 64))],
                                         args.cpu_seconds)
        cpu = {"value": round(flops / med_s / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
               "kind": kind,
               "sample": f"{reps} numeric-only multiplies of the same workload through the "
                         f"descriptor (fill done once, untimed), median of {reps}, one thread",
               "ms_per_multiply": round(med_s * 1e3, 3)}

    out = {
        "metric": "SpGEMM GFLOP/s (C=A*A numeric-only multiply, descriptor reuse, 1 GPU)",
        "value": round(flops / (ms_per_step * 1e-3) / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (5-point Laplacian, values U[0.5,1.5) redrawn per multiply)",
        "config": {"workload": f"laplacian5pt_g{args.grid}_AxA_numeric_reuse",
                   "rows": M, "nnz_a": A.nnz, "nnz_c": nnz_c, "products": products,
                   "algo": args.algo},
        "hbm_gbps": round(achieved, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic, "kernel_ms": round(avg_kernel_ms, 5),
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
        "phases_ms": {"work_estimation": round(t_sym * 1e3, 3),
                      "first_compute_fill_numeric": round(t_first * 1e3, 3)},
        "numeric_plan": [{k: p[k] for k in ("kind", "group", "cap", "rows")} for p in plan],
        "verified_bit_exact": verified,
    }
    print(json.dumps(out), flush=True)
    for p in val_bufs:
        rvsp.device_free(p)
    d.close()


if __name__ == "__main__":
    main()
