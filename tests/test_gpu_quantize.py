"""Encoder-side quantization on the GPU (pikgpu_quantize) against the
oracle's restatement (itself pinned bit for bit to the reference's
InitializeFrameEncCache + ComputeCoefficients): the int16 AC and DC bit for
bit, without predictions and with the LF / HF prediction residuals in every
combination, the decode of the device-resident result (the round trip) bit
for bit, and the refused cases."""
import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import PikGpuError, Reconstructor

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("w,h,seed,kw", [
    (13, 7, 2, {}), (136, 72, 4, dict(distance=2.0)),
    (83, 45, 22, dict(distance=0.5)), (264, 200, 11, dict(lut_ids="rand")),
    (600, 520, 3, {}), (1032, 776, 9, dict(distance=1.5)),
    (2048, 1536, 5, {})])
def test_matches_oracle(rec, w, h, seed, kw):
    from oracle import port
    print("quantize", w, h, seed, kw, flush=True)
    fr, op = synth.make_opsin_frame(w, h, seed=seed, predict=False, **kw)
    want_ac, want_dc = port.quantize(fr, op)
    rec.upload(fr)
    got_ac, got_dc = rec.quantize(op)
    assert np.array_equal(got_dc, want_dc)
    assert np.array_equal(got_ac, want_ac), int((got_ac != want_ac).sum())


@pytest.mark.parametrize("w,h,seed,kw", [
    (13, 7, 2, {}), (136, 72, 4, dict(distance=2.0)),
    (264, 200, 11, dict(lut_ids="rand")), (200, 136, 6, dict(lines=0.3)),
    (600, 520, 3, {}), (1032, 776, 9, dict(distance=1.5)),
    (2048, 1536, 5, {})])
@pytest.mark.parametrize("plf,phf", [(1, 1), (1, 0), (0, 1)])
def test_predictions_match_oracle(rec, w, h, seed, kw, plf, phf):
    from oracle import port
    print("quantize with predictions", w, h, seed, kw, plf, phf, flush=True)
    fr, op = synth.make_opsin_frame(w, h, seed=seed, predict=True, **kw)
    fr.predict_lf, fr.predict_hf = plf, phf
    want_ac, want_dc = port.quantize(fr, op)
    rec.upload(fr)
    got_ac, got_dc = rec.quantize(op)
    assert np.array_equal(got_dc, want_dc)
    assert np.array_equal(got_ac, want_ac), int((got_ac != want_ac).sum())


def test_predictions_against_reference(rec):
    # The default encode (both predictions on) directly against the
    # reference's ComputeCoefficients.
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    print("quantize with predictions vs reference", flush=True)
    fr, op = synth.make_opsin_frame(520, 264, seed=31, predict=True)
    want_ac, want_dc = ref.quantize(fr, op)
    rec.upload(fr)
    got_ac, got_dc = rec.quantize(op)
    assert np.array_equal(got_dc, want_dc)
    assert np.array_equal(got_ac, want_ac)


@pytest.mark.parametrize("predict", [False, True])
def test_round_trip_decode(rec, predict):
    # run() after quantize() decodes the quantized coefficients: the same
    # as uploading a frame that holds them.
    from oracle import port
    print("quantize round trip", predict, flush=True)
    fr, op = synth.make_opsin_frame(600, 520, seed=7, predict=predict)
    rec.upload(fr)
    ac, dc = rec.quantize(op)
    rec.run()
    got = rec.download()
    d = dict(vars(fr))
    d["ac"], d["dc"] = ac, dc
    rt = type(fr)(**d).contiguous()
    want, _ = port.recon(rt, stages=False)
    assert np.array_equal(_bits(got), _bits(want))


def test_refused_cases(rec):
    print("quantize refusals", flush=True)
    _, op = synth.make_opsin_frame(64, 64, seed=1)
    fr = synth.make_frame(64, 64, seed=1, num_dots=2)
    rec.upload(fr)
    with pytest.raises(PikGpuError) as e:
        rec.quantize(op)
    assert e.value.code == abi.EUNSUPPORTED
    fr, op = synth.make_opsin_frame(64, 64, seed=1, predict=False)
    rec.upload(fr)
    with pytest.raises(ValueError):
        rec.quantize(op[:, :8])
    # the context stays usable
    ac, dc = rec.quantize(op)
    assert ac.shape == (3, 8, 8 * 64) and dc.shape == (3, 8, 8)
