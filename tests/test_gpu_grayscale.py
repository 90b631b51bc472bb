"""Grayscale frames (FrameHeader::kGrayscaleOpt) on the GPU: the X and B DC
of every group rebuilt from Y (GrayXyb::RestoreXB, pik/compressed_image.cc:
94-209, 953-956) and the luma written to all three linear channels
(pik/pik_frame.cc:1155-1166), against the oracle's restatement (itself
pinned to the reference's own code) bit for bit: every stage at full size,
the downsampled decode, multi-pass frames, and the 8-bit output."""
import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import PikGpuError, Reconstructor

from .srgb import srgb8

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("w,h,seed,kw", [
    (13, 7, 2, {}), (136, 72, 4, dict(num_dots=3)),
    (200, 136, 21, dict(predict=False)), (600, 520, 3, dict(lut_ids="rand")),
    (1032, 776, 9, dict(num_dots=5))])
def test_stages_match_oracle(rec, w, h, seed, kw):
    from oracle import port
    print("grayscale", w, h, seed, kw, flush=True)
    fr = synth.make_frame(w, h, seed=seed, grayscale=True, **kw)
    want, wst = port.recon(fr, stages=True)
    rec.upload(fr)
    rec.run(keep_stages=True)
    got = rec.download()
    assert np.array_equal(_bits(got), _bits(want))
    assert np.array_equal(_bits(got[0]), _bits(got[2]))
    for s, name in enumerate(abi.STAGE_NAMES):
        assert np.array_equal(_bits(rec.stage(s)), _bits(wst[name])), name


def test_header_flag_spelling(rec):
    # FrameHeader::kGrayscaleOpt in the flag word decodes as grayscale = 1
    from oracle import port
    print("grayscale via header flags", flush=True)
    fr = synth.make_frame(200, 136, seed=14, grayscale=True)
    want, _ = port.recon(fr, stages=False)
    fr.grayscale = 0
    fr.flags = abi.FLAG_GRAYSCALE_OPT
    rec.upload(fr)
    rec.run()
    assert np.array_equal(_bits(rec.download()), _bits(want))


@pytest.mark.parametrize("ds", (2, 4, 8))
def test_downsampled_matches_oracle(rec, ds):
    from oracle import port
    print("grayscale downsampled", ds, flush=True)
    fr = synth.make_frame(520, 264, seed=5, grayscale=True)
    want = port.recon_downsampled(fr, ds)
    rec.set_downsample(ds)
    try:
        rec.upload(fr)
        rec.run()
        got = rec.download()
    finally:
        rec.set_downsample(1)
    assert np.array_equal(_bits(got), _bits(want))


def test_multipass_matches_oracle(rec):
    from oracle import port
    print("grayscale multipass", flush=True)
    fr = synth.make_frame(264, 200, seed=11, grayscale=True, num_dots=2)
    extra = synth.make_extra_passes(fr, 2, seed=4)
    want = port.recon_passes(fr, extra)
    rec.upload(fr)
    for e in extra:
        rec.add_pass(e)
    rec.run()
    assert np.array_equal(_bits(rec.download()), _bits(want))


def test_srgb8_gray_within_one_lsb(rec):
    from oracle import port
    print("grayscale srgb8", flush=True)
    fr = synth.make_frame(328, 200, seed=6, grayscale=True)
    want_lin, _ = port.recon(fr, stages=False)
    want = srgb8(want_lin).transpose(1, 2, 0)
    rec.upload(fr)
    rec.set_output(abi.OUTPUT_LINEAR | abi.OUTPUT_SRGB8)
    try:
        rec.run()
        got = rec.download_srgb8().astype(np.int32)
        lin = rec.download()
    finally:
        rec.set_output(abi.OUTPUT_LINEAR)
    assert np.array_equal(_bits(lin), _bits(want_lin))
    assert np.abs(got - want).max() <= 1
    assert np.array_equal(got[..., 0], got[..., 1])
    assert np.array_equal(got[..., 1], got[..., 2])


def test_forward_model_refuses_grayscale(rec):
    fr = synth.make_frame(64, 64, seed=1, grayscale=True)
    rec.upload(fr)
    qf = np.ones((8, 8), np.float32)
    with pytest.raises(PikGpuError):
        rec.fwd_adj(qf, target=np.zeros((3, 64, 64), np.float32))
