"""Pins the oracle's C restatement (oracle/pik_oracle.c) before it is trusted
as the checker:

  - against the committed golden fixtures (written from the reference's own
    code), bit for bit, for every stage the fixtures hold;
  - against the reference built from its sources (oracle/_ref), bit for bit
    at every stage, on frames that exercise each branch (all strategies,
    predictions on/off, every Gaborish strength, AR off, sharpened guide,
    both sigma LUTs, ragged sizes, several groups, dots);
  - its block transforms against the closed-form transposed-scaled DCT and
    round trips (mathematics, independent of the reference);
  - its EPF through invariants the reference's own tests use: a constant
    image is left unchanged by the whole filter + AR
    (pik/adaptive_reconstruction_test.cc:89-135 checks this invariance).
"""
import numpy as np
import pytest

from pik_amd import abi, synth

from .helpers import bit_equal_fraction, golden_names, load_golden, max_rel
from .test_gpu_parity import forced
from .test_host_transforms import size_of, slow_basis, slow_idct2


def _port():
    from oracle import port
    if not port.available():
        pytest.skip("oracle port not built (make -C oracle port)")
    return port


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_golden(name):
    port = _port()
    frame, want = load_golden(name)
    out, st = port.recon(frame)
    assert bit_equal_fraction(out, want["linear"]) == 1.0
    for key in ("dc_dequant", "dc", "idct", "ar"):
        assert bit_equal_fraction(st[key], want[key]) == 1.0, key


def _with(frame, **kw):
    for k, v in kw.items():
        setattr(frame, k, v)
    return frame


CASES = {
    "groups_lut_dots": lambda: synth.make_frame(600, 130, seed=2, num_dots=5,
                                                lut_ids="rand"),
    "ragged_nopredict": lambda: synth.make_frame(91, 67, seed=3,
                                                 predict=False),
    "single_block": lambda: synth.make_frame(8, 8, seed=4),
    "sub_block": lambda: synth.make_frame(5, 3, seed=5),
    "lf_only": lambda: _with(synth.make_frame(96, 64, seed=22),
                             predict_hf=0),
    "hf_only": lambda: _with(synth.make_frame(96, 64, seed=23),
                             predict_lf=0),
    "no_ar_g875": lambda: _with(synth.make_frame(96, 64, seed=24, num_dots=2),
                                adaptive_reconstruction=0,
                                gaborish=abi.GABORISH_875),
    "gab_off": lambda: _with(synth.make_frame(96, 64, seed=25),
                             gaborish=abi.GABORISH_OFF),
    "sharpened": lambda: _with(synth.make_frame(96, 64, seed=26),
                               epf_use_sharpened=1),
    "d3": lambda: synth.make_frame(128, 96, seed=27, distance=3.0),
    # LINES blocks among the others (the reference decodes them; its encoder
    # never chooses them), with and without predictions
    "lines_mixed": lambda: synth.make_frame(136, 104, seed=28, num_dots=2,
                                            lines=0.3, lut_ids="rand"),
    "lines_nopredict": lambda: synth.make_frame(96, 72, seed=29, lines=0.4,
                                                predict=False),
}
for _raw in (abi.ACS_IDENTITY, abi.ACS_DCT2X2, abi.ACS_DCT4X4,
             abi.ACS_DCT16X16, abi.ACS_DCT32X32, abi.ACS_DCT_NOHF,
             abi.ACS_DCT4X4_NOHF, abi.ACS_LINES):
    CASES["all_raw%d" % _raw] = (lambda raw=_raw: forced(synth.make_frame(
        64, 64, seed=40 + raw, lut_ids="rand"), raw))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_at_every_stage(name):
    port = _port()
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    frame = CASES[name]()
    out_r, st_r = ref.recon(frame, threads=1)
    out_p, st_p = port.recon(frame)
    for s, sname in enumerate(abi.STAGE_NAMES):
        if not frame.adaptive_reconstruction and s in (abi.STAGE_GUIDE,
                                                       abi.STAGE_EPF):
            continue
        assert bit_equal_fraction(st_p[sname], st_r[sname]) == 1.0, (
            sname, max_rel(st_p[sname], st_r[sname]))
    assert bit_equal_fraction(out_p, out_r) == 1.0


@pytest.mark.parametrize("raw", [abi.ACS_DCT, abi.ACS_DCT16X16,
                                 abi.ACS_DCT32X32])
def test_oracle_transforms_closed_form(raw):
    port = _port()
    n = size_of(raw)
    rng = np.random.default_rng(7 + raw)
    coef = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    got = port.transform(raw, 0, coef)
    assert np.abs(got - slow_idct2(coef)).max() < 2e-5
    pix = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    b = slow_basis(n)
    want = np.linalg.solve(b.T, np.linalg.solve(b.T, pix.astype(np.float64)).T)
    assert np.abs(port.transform(raw, 1, pix) - want).max() < 2e-6


@pytest.mark.parametrize("raw", [abi.ACS_DCT, abi.ACS_IDENTITY,
                                 abi.ACS_DCT2X2, abi.ACS_DCT4X4,
                                 abi.ACS_DCT16X16, abi.ACS_DCT32X32,
                                 abi.ACS_LINES])
def test_oracle_transform_round_trip_and_reference(raw):
    port = _port()
    n = size_of(raw)
    rng = np.random.default_rng(70 + raw)
    pix = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    back = port.transform(raw, 0, port.transform(raw, 1, pix))
    keep = np.ones((n, n), bool)
    if raw == abi.ACS_LINES:
        # the four pixels on no line are interpolated, not reproduced
        keep.reshape(-1)[[21, 28, 35, 42]] = False
    assert np.abs(back - pix)[keep].max() < 2e-5
    from oracle import ref
    if ref.available("scalar"):
        for d in (0, 1):
            blk = rng.uniform(-1, 1, (n, n)).astype(np.float32)
            assert np.array_equal(port.transform(raw, d, blk).view(np.uint32),
                                  ref.transform(raw, d, blk).view(np.uint32))


def test_constant_image_is_a_fixed_point_of_epf_and_ar():
    """A frame of DCT8 blocks with zero AC and one DC value decodes to a flat
    XYB image; EPF then averages equal values and AR sees no correction, so
    the AR stage equals the IDCT stage up to the rounding of a 56-term
    float sum (pik/adaptive_reconstruction_test.cc:89-135 checks the same
    invariance)."""
    port = _port()
    frame = synth.make_frame(64, 48, seed=9)
    frame.ac[:] = 0
    frame.ac_strategy[:] = abi.ACS_DCT << 4
    frame.predict_lf = frame.predict_hf = 0
    for c in range(3):
        frame.dc[c][:] = frame.dc[c].flat[0]
    frame.dots = frame.dots[:0]
    _, st = port.recon(frame)
    for c in range(3):
        assert np.ptp(st["idct"][c]) == 0.0
    assert max_rel(st["ar"], st["idct"]) < 1e-6
