"""Writes tests/golden/search_fields_1024.npz: the reference encoder's
initial quant field (InitialQuantField, pik/adaptive_quantization.cc:
1047-1060, at Butteraugli distance 1.0) of each of the first 16 seeded
1024x1024 search images of pik_amd.synth.make_search_inputs, and
InitialQuantDC (:1040-1045). The quant-field search benchmark starts the GPU
search and the reference's own search from these. The reference code is the
scalar build under oracle/_ref (make -C oracle ref); run from the repository
root:

    python3 tests/golden/make_search_fields.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pik_amd import synth  # noqa: E402

SIZE = 1024
SEEDS = range(16)
TARGET = 1.0
NUM_RECTS = 40


def main():
    fields = []
    for seed in SEEDS:
        _, _, orig, _ = synth.make_search_inputs(SIZE, SIZE, seed=seed,
                                                 num_rects=NUM_RECTS,
                                                 target=TARGET)
        fields.append(ref.initial_quant_field(orig, TARGET))
        print("seed", seed, flush=True)
    path = os.path.join(HERE, "search_fields_%d.npz" % SIZE)
    np.savez_compressed(path, fields=np.stack(fields).astype(np.float32),
                        seeds=np.array(list(SEEDS)), size=SIZE,
                        target=TARGET, num_rects=NUM_RECTS,
                        quant_dc=ref.initial_quant_dc(TARGET))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
