"""Writes tests/golden/dequant_default.npz: the reference's library
dequantisation tables (DequantMatrices::Compute, pik/quant_weights.cc:
607-926) as produced by the reference build under oracle/_ref, the fixture
the library's own computation is checked against. Run from the repo root
after `make -C oracle ref`."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))
from oracle import ref as refbind  # noqa: E402

t, o = refbind.default_tables()
out = os.path.join(refbind.ROOT, "tests", "golden", "dequant_default.npz")
np.savez_compressed(out, table=t, offsets=o)
print("wrote", out, t.shape, o.shape, o)
