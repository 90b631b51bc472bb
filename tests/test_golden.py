"""The committed golden fixtures (tests/golden/recon_*.npz, written by
tests/golden/make_golden.py from the reference's own code) against the
oracles, on the CPU."""
import numpy as np
import pytest

from pik_amd import abi
from pik_amd.runtime import validate

from .helpers import bit_equal_fraction, golden_names, load_golden


def test_fixtures_present():
    assert len(golden_names()) >= 5


@pytest.mark.parametrize("name", golden_names())
def test_fixture_is_decodable(name):
    frame, want = load_golden(name)
    validate(frame)
    assert want["linear"].shape == (3, frame.ysize, frame.xsize)
    assert want["dc"].shape == (3, frame.ysize_blocks, frame.xsize_blocks)
    assert all(np.isfinite(v).all() for v in want.values())


@pytest.mark.parametrize("name", golden_names())
def test_reference_reproduces_fixture(name):
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    frame, want = load_golden(name)
    out, st = ref.recon(frame, threads=2, stages=True)
    assert bit_equal_fraction(out, want["linear"]) == 1.0
    for key in ("dc_dequant", "dc", "idct", "ar"):
        assert bit_equal_fraction(st[key], want[key]) == 1.0, key
