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
    ev0 = [hip.event() for _ in range(steps)]
    ev1 = [hip.event() for _ in range(steps)]
    hip.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        hip.record(ev0[i], stream)
        st = step(i)
        hip.record(ev1[i], stream)
        if st != 0:
            F.check(st, "compute")
    hip.sync()
    wall = time.perf_counter() - t0
    kernel_ms = [hip.elapsed_ms(ev0[i], ev1[i]) for i in range(steps)]
    plan = [p for p in d.plan() if p["phase"] == F.PHASE_NUMERIC and p["rows"] > 0]
    d.close()
    return {"ms_per_step": wall * 1e3 / steps, "kernel_ms": statistics.mean(kernel_ms),
            "kernel_ms_min": min(kernel_ms), "t_sym": t_sym, "t_first": t_first,
            "t_tune": t_tune, "plan": plan, "nnz_c": nnz_c, "C": dC}


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
        print("bench.py measures one GPU: every box in this setup has one MI355X, and "
              "the SpGEMM path has no multi-GPU mode (DESIGN.md describes how it would "
              "shard by output rows).", file=sys.stderr)
        sys.exit(2)

    import rvsp
    from rvsp import _ffi as F
    from rvsp import gen

    hip = Hip()
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
    }
    print(json.dumps(out), flush=True)
    for r in runs.values():
        r["C"].free()
    for p in val_bufs:
        rvsp.device_free(p)
    dA.free()


if __name__ == "__main__":
    main()
