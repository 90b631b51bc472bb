"""The library dequantisation tables (DequantMatrices::Compute for the
predefined table set, pik/quant_weights.cc:607-926), computed by the host
side of libpikgpu.so (pikgpu_default_dequant_tables)."""
import ctypes

import numpy as np

from . import abi

_cache = None


def default_tables():
    """Returns (table float32[4800], offsets uint32[21])."""
    global _cache
    if _cache is None:
        lib = abi.load()
        ts, no = ctypes.c_size_t(), ctypes.c_size_t()
        lib.pikgpu_default_dequant_tables(None, 0, None, 0, ctypes.byref(ts),
                                          ctypes.byref(no))
        table = np.zeros(ts.value, np.float32)
        offsets = np.zeros(no.value, np.uint32)
        lib.pikgpu_default_dequant_tables(table.ctypes.data, table.size,
                                          offsets.ctypes.data, offsets.size,
                                          ctypes.byref(ts), ctypes.byref(no))
        _cache = (table, offsets)
    return _cache
