"""The oracle on frames whose quant field, colour correlation and quantized
coefficients come from the reference encoder (tests/refside.py), against
the reference decoder at every stage, on the CPU."""
import numpy as np
import pytest

from pik_amd import abi

from .helpers import bit_equal_fraction


def _need_ref():
    from oracle import ref, port
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    if not port.available():
        pytest.skip("oracle not built")
    return ref, port


@pytest.mark.parametrize("w,h,seed,predict,d", [
    (136, 104, 1, True, 1.0), (200, 136, 2, False, 1.0),
    (96, 72, 3, True, 2.5)])
def test_oracle_on_reference_side_information(w, h, seed, predict, d):
    ref, port = _need_ref()
    from .refside import reference_side_frame
    fr = reference_side_frame(w, h, seed, distance=d, predict=predict)
    out_r, st_r = ref.recon(fr, threads=1)
    out_p, st_p = port.recon(fr)
    for name in abi.STAGE_NAMES:
        assert bit_equal_fraction(st_p[name], st_r[name]) == 1.0, name
    assert bit_equal_fraction(out_p, out_r) == 1.0
    # the side information is the reference's, not the generator's
    assert fr.ytox_map.dtype == np.int32 and fr.global_scale not in (8192,)
