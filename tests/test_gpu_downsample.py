"""Downsampled decode on the GPU (pikgpu_set_downsample, factors 2, 4, 8)
against the fixtures written from the reference's own code and against the
oracle's restatement on seeded frames: bit for bit. Also the 8-bit output at
the downsampled size, the refused cases (dots, stage dumps, bad factors) and
the return to full size."""
import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import PikGpuError, Reconstructor

from .helpers import downsample_names, load_downsample

pytestmark = pytest.mark.gpu

FACTORS = (2, 4, 8)


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _decode(rec, frame, ds):
    rec.set_downsample(ds)
    try:
        rec.upload(frame)
        rec.run()
        return rec.download()
    finally:
        rec.set_downsample(1)


@pytest.mark.parametrize("name", downsample_names())
@pytest.mark.parametrize("ds", FACTORS)
def test_fixture(rec, name, ds):
    print("fixture", name, "factor", ds, flush=True)
    frame, want = load_downsample(name)
    got = _decode(rec, frame, ds)
    assert got.shape == want[ds].shape
    assert np.array_equal(_bits(got), _bits(want[ds]))


@pytest.mark.parametrize("w,h,seed", [(13, 7, 2), (130, 75, 7), (600, 520, 3),
                                      (1032, 776, 9)])
@pytest.mark.parametrize("ds", FACTORS)
def test_matches_oracle(rec, w, h, seed, ds):
    from oracle import port
    print("seeded", w, h, seed, "factor", ds, flush=True)
    fr = synth.make_frame(w, h, seed=seed, num_dots=0)
    want = port.recon_downsampled(fr, ds)
    got = _decode(rec, fr, ds)
    assert np.array_equal(_bits(got), _bits(want))


def test_srgb8_and_return_to_full_size(rec):
    from oracle import port
    from .srgb import srgb8
    fr = synth.make_frame(520, 400, seed=31, num_dots=0)
    lin4 = port.recon_downsampled(fr, 4)
    rec.upload(fr)
    rec.set_downsample(4)
    rec.set_output(abi.OUTPUT_LINEAR | abi.OUTPUT_SRGB8)
    try:
        rec.run()
        got = rec.download_srgb8().astype(np.int32)
        assert np.array_equal(_bits(rec.download()), _bits(lin4))
    finally:
        rec.set_output(abi.OUTPUT_LINEAR)
        rec.set_downsample(1)
    want = srgb8(lin4).transpose(1, 2, 0)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1
    # full size again on the same upload: the ordinary decode
    rec.run()
    full = rec.download()
    assert full.shape == (3, fr.ysize, fr.xsize)
    ref_full, _ = port.recon(fr, stages=False)
    assert np.array_equal(_bits(full), _bits(ref_full))


def test_refused_cases(rec):
    fr = synth.make_frame(64, 64, seed=1, num_dots=2)
    rec.upload(fr)
    with pytest.raises(PikGpuError) as e:
        rec.set_downsample(3)
    assert e.value.code == abi.EINVAL
    rec.set_downsample(2)
    try:
        with pytest.raises(PikGpuError) as e:
            rec.run()  # dots cannot be downsampled
        assert e.value.code == abi.EUNSUPPORTED
        fr0 = synth.make_frame(64, 64, seed=1, num_dots=0)
        rec.upload(fr0)
        with pytest.raises(PikGpuError) as e:
            rec.run(keep_stages=True)
        assert e.value.code == abi.EINVAL
        rec.run()
        with pytest.raises(PikGpuError) as e:
            rec.stage(abi.STAGE_IDCT)  # the last run was downsampled
        assert e.value.code == abi.ESTATE
    finally:
        rec.set_downsample(1)
