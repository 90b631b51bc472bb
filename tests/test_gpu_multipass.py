"""Multi-pass accumulation on the GPU (pikgpu_frame_add_pass) against the
oracle's restatement (itself pinned to the reference's own code), bit for
bit; at 4096 x 4096 the size-independent property that splitting the AC
over disjoint passes reproduces the single-pass decode bit for bit; and
the refused cases."""
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


def _decode(rec, frame, extra, ds=1):
    rec.upload(frame)
    for e in extra:
        rec.add_pass(e)
    rec.set_downsample(ds)
    try:
        rec.run()
        return rec.download()
    finally:
        rec.set_downsample(1)


@pytest.mark.parametrize("w,h,seed,npass,dots", [
    (13, 7, 2, 1, 0), (130, 75, 7, 1, 2), (264, 200, 11, 2, 3),
    (600, 520, 5, 3, 4), (1032, 776, 9, 2, 5)])
def test_matches_oracle(rec, w, h, seed, npass, dots):
    from oracle import port
    print("passes", w, h, seed, npass, flush=True)
    fr = synth.make_frame(w, h, seed=seed, num_dots=dots)
    extra = synth.make_extra_passes(fr, npass, seed=seed)
    want = port.recon_passes(fr, extra)
    assert np.array_equal(_bits(_decode(rec, fr, extra)), _bits(want))


def test_disjoint_split_at_4096(rec):
    print("split 4096", flush=True)
    fr = synth.make_frame(4096, 4096, seed=2, num_dots=16)
    single = _decode(rec, fr, [])
    first, extra = synth.split_passes(fr, 3, seed=5)
    assert np.array_equal(_bits(_decode(rec, first, extra)), _bits(single))


def test_split_with_downsampling(rec):
    fr = synth.make_frame(520, 400, seed=6, num_dots=0)
    single = _decode(rec, fr, [], ds=4)
    first, extra = synth.split_passes(fr, 2, seed=1)
    assert np.array_equal(_bits(_decode(rec, first, extra, ds=4)),
                          _bits(single))


def test_refused_cases(rec):
    fr = synth.make_frame(64, 64, seed=1, num_dots=0)
    extra = synth.make_extra_passes(fr, 1, seed=1)
    with Reconstructor(0) as fresh:
        with pytest.raises(PikGpuError) as e:
            fresh.add_pass(extra[0])  # nothing uploaded
        assert e.value.code == abi.ESTATE
    rec.upload(fr)
    for _ in range(15):
        rec.add_pass(extra[0])
    with pytest.raises(PikGpuError) as e:
        rec.add_pass(extra[0])  # a 17th pass
    assert e.value.code == abi.EINVAL
    with pytest.raises(PikGpuError) as e:
        rec.fwd_adj(fr.quant_field.astype(np.float32),
                    target=np.zeros((3, fr.ysize, fr.xsize), np.float32))
    assert e.value.code == abi.EUNSUPPORTED
    # a new upload starts again from one pass
    from oracle import port
    want = port.recon_passes(fr, [])
    assert np.array_equal(_bits(_decode(rec, fr, [])), _bits(want))
