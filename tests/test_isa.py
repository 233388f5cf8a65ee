"""Static checks of the gfx950 code the library is built from.

Bit parity with the reference needs a separately rounded multiply and add:
no fused multiply-add may appear in the device code (hipcc contracts a*b+c
into v_fma_f32 by default). Packed FP32 instructions are switched off for
all device code. Both are checked on the device assembly hipcc produces from
the library sources with the library's own flags.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rv-sparse_amd")
HIPCC = "/opt/rocm/bin/hipcc"


def device_asm(src):
    flags = subprocess.run(["make", "-s", "-C", PKG, "print-hipflags"], capture_output=True,
                           text=True, check=True).stdout.split()
    out = subprocess.run([HIPCC, *flags, "--cuda-device-only", "-S", "-o", "-", src],
                         cwd=PKG, capture_output=True, text=True, check=True)
    return out.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", sorted(f for f in os.listdir(os.path.join(PKG, "csrc"))
                                       if f.endswith(".hip")))
def test_no_fma_no_packed_fp32(src):
    asm = device_asm(os.path.join("csrc", src))
    assert "v_add_f32" in asm or "v_mul_f32" in asm or "kernels" not in src
    fused = re.findall(r"\b(v_fma_f(?:32|64)|v_fmac_f(?:32|64)\w*|v_fma_mix\w*|v_mad_f32|"
                       r"v_mac_f32)\b", asm)
    assert not fused, f"fused multiply-add in device code: {sorted(set(fused))}"
    packed = re.findall(r"\bv_pk_(?:add|mul|fma)_f32\b", asm)
    assert not packed, f"packed FP32 instructions in device code: {sorted(set(packed))}"
