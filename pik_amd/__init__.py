"""MI355X (gfx950) implementation of the PIK decoder-side reconstruction.

The product is the HIP library libpikgpu.so behind the C ABI in
include/pikgpu.h; this package holds its sources (csrc/), the ctypes
description of the ABI (abi.py), the frame container (frame.py), the
host-side runtime wrapper (runtime.py) and the synthetic input generator
(synth.py).
"""
