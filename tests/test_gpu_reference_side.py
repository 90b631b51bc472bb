"""GPU reconstruction of frames whose quant field, colour correlation and
quantized coefficients come from the reference encoder (InitialQuantField,
FindBestColorCorrelationMap, the reference's quantization; tests/refside.py)
against the reference decoder: bit for bit at every stage."""
import pytest

from pik_amd import abi
from pik_amd.runtime import Reconstructor

from .helpers import bit_equal_fraction, max_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


@pytest.mark.parametrize("w,h,seed,predict,d", [
    (600, 520, 11, True, 1.0), (1032, 776, 12, True, 1.0),
    (333, 257, 13, False, 2.0), (512, 512, 14, True, 0.6)])
def test_stages_bit_exact(rec, w, h, seed, predict, d):
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    from .refside import reference_side_frame
    print("reference side information", w, h, seed, predict, d, flush=True)
    fr = reference_side_frame(w, h, seed, distance=d, predict=predict)
    want, st = ref.recon(fr, threads=8)
    rec.upload(fr)
    rec.run(keep_stages=True)
    for s, name in enumerate(abi.STAGE_NAMES):
        got = rec.stage(s)
        assert max_rel(got, st[name]) <= 1e-4, name
        assert bit_equal_fraction(got, st[name]) == 1.0, name
    assert bit_equal_fraction(rec.download(), want) == 1.0
