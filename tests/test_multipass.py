"""Multi-pass accumulation (AC passes after the frame's own, Dequant<false>)
on the CPU: the oracle's restatement against the reference's own code
(harness: DequantImageAC per pass added into the first pass's
coefficients, as Dequant<false>::DoAC does), bit for bit, and the
restatement's split invariance."""
import numpy as np
import pytest

from pik_amd import synth


def _port():
    from oracle import port
    if not port.available():
        pytest.skip("oracle port not built (make -C oracle port)")
    return port


def _ref():
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    return ref


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("w,h,seed,npass,dots", [
    (13, 7, 2, 1, 0), (130, 75, 7, 1, 2), (264, 200, 11, 2, 3),
    (600, 520, 5, 3, 4)])
def test_port_matches_reference(w, h, seed, npass, dots):
    ref, port = _ref(), _port()
    fr = synth.make_frame(w, h, seed=seed, num_dots=dots)
    extra = synth.make_extra_passes(fr, npass, seed=seed)
    want = ref.recon_passes(fr, extra, threads=2)
    got = port.recon_passes(fr, extra)
    assert np.array_equal(_bits(got), _bits(want))
    # the passes change the image
    assert not np.array_equal(_bits(got), _bits(port.recon_passes(fr, [])))


def test_zero_pass_is_identity():
    port = _port()
    fr = synth.make_frame(96, 64, seed=3, num_dots=1)
    zero = [np.zeros_like(fr.ac)]
    assert np.array_equal(_bits(port.recon_passes(fr, zero)),
                          _bits(port.recon_passes(fr, [])))


@pytest.mark.parametrize("npass", [2, 3])
def test_disjoint_split_equals_single_pass(npass):
    port = _port()
    fr = synth.make_frame(200, 136, seed=8, num_dots=2)
    first, extra = synth.split_passes(fr, npass, seed=npass)
    assert not np.array_equal(first.ac, fr.ac)
    assert np.array_equal(_bits(port.recon_passes(first, extra)),
                          _bits(port.recon_passes(fr, [])))
