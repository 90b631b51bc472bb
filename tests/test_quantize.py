"""Encoder-side quantization (the coefficient half of RoundtripImage,
pik/adaptive_quantization.cc:580-651) on the CPU: the oracle's C
restatement of InitializeFrameEncCache + ComputeCoefficients
(pik/compressed_image.cc:210-455) against the reference's own functions,
called by the harness as RoundtripImage calls them, bit for bit on seeded
frames without predictions and with the LF / HF prediction residuals
(PredictLfForEncoder + ComputePredictionResiduals, pik/ac_predictions.cc:
477-557, 637-650) in every combination; and the round trip itself:
decoding the quantized frame gives back the image it was made from, to
within the quantization error."""
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


CASES = [(13, 7, 2, {}), (64, 64, 1, {}), (136, 72, 4, dict(distance=2.0)),
         (83, 45, 22, dict(distance=0.5)),
         (264, 200, 11, dict(lut_ids="rand")), (600, 520, 3, {}),
         (1032, 776, 9, dict(distance=1.5))]


@pytest.mark.parametrize("w,h,seed,kw", CASES)
def test_port_matches_reference(w, h, seed, kw):
    ref, port = _ref(), _port()
    fr, op = synth.make_opsin_frame(w, h, seed=seed, predict=False, **kw)
    want_ac, want_dc = ref.quantize(fr, op)
    got_ac, got_dc = port.quantize(fr, op)
    assert np.array_equal(got_dc, want_dc)
    assert np.array_equal(got_ac, want_ac), int((got_ac != want_ac).sum())


PRED_CASES = [(13, 7, 2, {}), (64, 64, 1, {}), (136, 72, 4, dict(distance=2.0)),
              (264, 200, 11, dict(lut_ids="rand")),
              (200, 136, 6, dict(lines=0.3)), (600, 520, 3, {}),
              (1032, 776, 9, dict(distance=1.5))]


@pytest.mark.parametrize("w,h,seed,kw", PRED_CASES)
@pytest.mark.parametrize("plf,phf", [(1, 1), (1, 0), (0, 1)])
def test_port_matches_reference_with_predictions(w, h, seed, kw, plf, phf):
    ref, port = _ref(), _port()
    fr, op = synth.make_opsin_frame(w, h, seed=seed, predict=True, **kw)
    fr.predict_lf, fr.predict_hf = plf, phf
    want_ac, want_dc = ref.quantize(fr, op)
    got_ac, got_dc = port.quantize(fr, op)
    assert np.array_equal(got_dc, want_dc)
    assert np.array_equal(got_ac, want_ac), int((got_ac != want_ac).sum())


def test_predictions_change_the_coefficients():
    # The residuals are not the plain coefficients: with predictions on, the
    # low- and high-frequency positions of most blocks differ.
    port = _port()
    fr, op = synth.make_opsin_frame(200, 136, seed=7, predict=True)
    with_pred, _ = port.quantize(fr, op)
    fr.predict_lf = fr.predict_hf = 0
    plain, _ = port.quantize(fr, op)
    assert (with_pred != plain).mean() > 0.01


def _with_coefficients(frame, ac, dc):
    d = dict(vars(frame))
    d["ac"], d["dc"] = ac, dc
    return type(frame)(**d).contiguous()


@pytest.mark.parametrize("w,h,seed,predict", [(136, 72, 4, False),
                                              (264, 200, 11, False),
                                              (264, 200, 11, True)])
def test_round_trip_reproduces_image(w, h, seed, predict):
    # The decode of the quantized frame against the image it came from, both
    # in linear RGB: at distance 1 the mean error is a small fraction of the
    # image's own spread (3 and 1.7 of about 25 on these frames), where a
    # swapped channel, a sign or a scale mistake is of the spread's size.
    port = _port()
    fr, op = synth.make_opsin_frame(w, h, seed=seed, predict=predict)
    ac, dc = port.quantize(fr, op)
    rt = _with_coefficients(fr, ac, dc)
    rt.adaptive_reconstruction = 0
    rt.gaborish = 0
    rt.dots = rt.dots[:0]
    out, _ = port.recon(rt, stages=False)
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    want = ref.opsin_to_linear(op)[:, :h, :w]
    err = np.abs(out.astype(np.float64) - want)
    spread = np.abs(want - want.mean(axis=(1, 2), keepdims=True)).mean()
    assert err.mean() < 0.15 * spread, (err.mean(), spread)
    assert np.percentile(err, 99) < 25.0


def test_refuses_dots_and_grayscale():
    port = _port()
    fr, op = synth.make_opsin_frame(64, 64, seed=1, predict=True)
    fr.grayscale = 1
    with pytest.raises(RuntimeError, match="grayscale"):
        port.quantize(fr, op)


def test_zero_image_quantizes_to_bias_only():
    # A constant image has no AC energy: every AC coefficient quantizes to
    # zero (positions k > 0 of every sub-block), whatever the strategy.
    ref, port = _ref(), _port()
    fr, op = synth.make_opsin_frame(128, 64, seed=5, predict=False)
    flat = np.empty_like(op)
    for c in range(3):
        flat[c] = op[c].mean()
    want_ac, want_dc = ref.quantize(fr, flat)
    got_ac, got_dc = port.quantize(fr, flat)
    assert np.array_equal(got_ac, want_ac) and np.array_equal(got_dc, want_dc)
    # only the LLF positions can be non-zero: one per 8x8 sub-block
    assert int((got_ac != 0).sum()) <= 3 * fr.xsize_blocks * fr.ysize_blocks
