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
