"""The library dequantisation tables computed by libpikgpu
(pikgpu_default_dequant_tables, a restatement of DequantMatrices::Compute,
pik/quant_weights.cc:416-566, 607-926) against the reference's own tables:
the committed fixture tests/golden/dequant_default.npz (written by
tests/golden/make_tables.py from oracle/_ref) and, where it is built, the
reference build itself."""
import os

import numpy as np

from pik_amd import tables

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "dequant_default.npz")
KIND_SIZES = [64, 64, 64, 64, 256, 1024, 64]


def test_layout():
    t, o = tables.default_tables()
    assert t.size == 3 * sum(KIND_SIZES) == 4800
    expect, pos = [], 0
    for n in KIND_SIZES:
        for _ in range(3):
            expect.append(pos)
            pos += n
    assert list(o) == expect


def test_bit_exact_against_fixture():
    t, o = tables.default_tables()
    z = np.load(GOLDEN, allow_pickle=False)
    assert np.array_equal(o, z["offsets"])
    assert np.array_equal(t.view(np.uint32), z["table"].view(np.uint32))


def test_bit_exact_against_reference_build():
    from oracle import ref
    if not ref.available("scalar"):
        import pytest
        pytest.skip("reference harness not built (oracle/_ref)")
    t, o = tables.default_tables()
    rt, ro = ref.default_tables()
    assert np.array_equal(o, ro)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))


def test_values_are_positive_and_finite():
    t, _ = tables.default_tables()
    assert np.all(np.isfinite(t)) and np.all(t > 0)
