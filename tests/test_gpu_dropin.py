"""The reference's own C test programs and examples (tests/test_*.c,
examples/*.c), compiled from its sources against this project's headers and
librvsparse_hip.so (oracle/Makefile `dropin`), must pass and print exactly
what the same sources print when linked against the reference library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")
REFBIN = os.path.join(ROOT, "oracle", "_ref", "refbin")
PROGRAMS = sorted(os.listdir(DROPIN)) if os.path.isdir(DROPIN) else []


@pytest.mark.skipif(not PROGRAMS, reason="reference programs not built (no reference tree at build)")
@pytest.mark.parametrize("prog", PROGRAMS)
def test_reference_program_on_gpu_library(gpu, prog):
    ours = subprocess.run([os.path.join(DROPIN, prog)], capture_output=True, text=True,
                          timeout=120)
    theirs = subprocess.run([os.path.join(REFBIN, prog)], capture_output=True, text=True,
                            timeout=120)
    print(ours.stdout[-2000:])
    assert ours.returncode == 0, ours.stdout + ours.stderr
    assert ours.stdout == theirs.stdout


BENCH = os.path.join(ROOT, "oracle", "_ref", "bench")


@pytest.mark.skipif(not os.path.exists(os.path.join(BENCH, "dropin_bench")),
                    reason="reference bench harness not built (no reference tree at build)")
@pytest.mark.parametrize("gen", [("1500", "1500", "0.004", "3"), ("3000", "700", "0.01", "5")])
def test_reference_bench_harness_on_gpu_library(gpu, gen):
    """bench/bench.c of the reference, unchanged, calling the raw scalar kernel
    that this library serves on the GPU: same matrix statistics (nnz of C,
    flops, per-row product counts) as the reference library, and its own
    correctness column passes."""
    args = ["--kernel", "scalar_f32", "--gen", *gen, "--runs", "2", "--warmup", "1"]
    ours = subprocess.run([os.path.join(BENCH, "dropin_bench"), *args], capture_output=True,
                          text=True, timeout=120)
    theirs = subprocess.run([os.path.join(BENCH, "ref_bench"), *args], capture_output=True,
                            text=True, timeout=120)
    print(ours.stdout[-2000:])
    assert ours.returncode == 0, ours.stdout + ours.stderr
    rows_o = [l.split(",") for l in ours.stdout.splitlines() if l.startswith("-,scalar_f32")]
    rows_t = [l.split(",") for l in theirs.stdout.splitlines() if l.startswith("-,scalar_f32")]
    assert len(rows_o) == len(rows_t) == 2
    for o, t in zip(rows_o, rows_t):
        assert o[8:17] == t[8:17]   # rows .. op_var
        assert o[20] == "1"          # correct
