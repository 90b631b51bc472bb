"""Dequantisation tables of the reference's library encodings
(DequantMatrices::Compute, pik/quant_weights.cc:607-926).

The table (4800 floats) and its offsets are data of the format: a decoder
uses them to turn quantized coefficients into values. They are loaded from
pik_amd/data/dequant_default.npz.
"""
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                     "dequant_default.npz")
_cache = None


def default_tables():
    """Returns (table float32[4800], offsets uint32[21])."""
    global _cache
    if _cache is None:
        z = np.load(_PATH, allow_pickle=False)
        _cache = (np.ascontiguousarray(z["table"], np.float32),
                  np.ascontiguousarray(z["offsets"], np.uint32))
    return _cache
