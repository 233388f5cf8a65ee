"""rvsp — Python host mirror of the rv-sparse C API served by the gfx950 HIP
library librvsparse_hip.so (built from rv-sparse_amd/csrc)."""
from . import _ffi
from ._ffi import RvspError, lib
from .api import (Csr, DeviceCsr, Descriptor, device_alloc, device_count,
                  device_free, from_device, spgemm, spgemm_device, spgemm_legacy,
                  spgemm_oneshot, synchronize, to_device)

__all__ = ["_ffi", "RvspError", "lib", "Csr", "DeviceCsr", "Descriptor",
           "device_alloc", "device_count", "device_free", "from_device",
           "spgemm", "spgemm_device", "spgemm_legacy", "spgemm_oneshot", "synchronize",
           "to_device"]
