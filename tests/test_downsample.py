"""Downsampled decode (factors 2, 4, 8) on the CPU: the committed fixtures
(tests/golden/downsample_*.npz, written from the reference's own
ReconOpsinImage / FinalizeFrameDecoding with `downsample`) against the
reference harness and the oracle's C restatement, and the restatement
against the reference on seeded frames, bit for bit."""
import numpy as np
import pytest

from pik_amd import synth

from .helpers import downsample_names, load_downsample

FACTORS = (2, 4, 8)


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


def test_fixtures_present():
    assert len(downsample_names()) >= 3


@pytest.mark.parametrize("name", downsample_names())
def test_fixture_shapes(name):
    frame, want = load_downsample(name)
    assert len(frame.dots) == 0
    for ds in FACTORS:
        assert want[ds].shape == (3, -(-frame.ysize // ds),
                                  -(-frame.xsize // ds))
        assert np.isfinite(want[ds]).all()


@pytest.mark.parametrize("name", downsample_names())
def test_reference_reproduces_fixture(name):
    ref = _ref()
    frame, want = load_downsample(name)
    for ds in FACTORS:
        assert np.array_equal(_bits(ref.recon_downsampled(frame, ds)),
                              _bits(want[ds])), ds


@pytest.mark.parametrize("name", downsample_names())
def test_port_matches_fixture(name):
    port = _port()
    frame, want = load_downsample(name)
    for ds in FACTORS:
        assert np.array_equal(_bits(port.recon_downsampled(frame, ds)),
                              _bits(want[ds])), ds


@pytest.mark.parametrize("w,h,seed", [(13, 7, 2), (64, 64, 5), (264, 200, 11),
                                      (600, 520, 3)])
def test_port_matches_reference(w, h, seed):
    ref, port = _ref(), _port()
    fr = synth.make_frame(w, h, seed=seed, num_dots=0)
    for ds in FACTORS:
        assert np.array_equal(_bits(port.recon_downsampled(fr, ds)),
                              _bits(ref.recon_downsampled(fr, ds))), ds


def test_flat_frame_is_flat():
    # A frame whose AC coefficients are all zero and whose DC is constant
    # decodes to a constant image at every factor; the mean of a constant
    # is that constant exactly (sums of ds^2 equal floats scaled by the
    # power of two 1/ds^2).
    port = _port()
    fr = synth.make_frame(64, 48, seed=1, num_dots=0)
    for c in range(3):
        fr.ac[c][...] = 0
        fr.dc[c][...] = 7 if c == 1 else 0
    fr.adaptive_reconstruction = 0
    full_ds = {ds: port.recon_downsampled(fr, ds) for ds in FACTORS}
    for ds, img in full_ds.items():
        for c in range(3):
            assert np.ptp(img[c]) == 0.0, (ds, c)
    assert np.array_equal(full_ds[2][:, 0, 0], full_ds[8][:, 0, 0])


def test_dots_are_refused():
    port = _port()
    fr = synth.make_frame(64, 64, seed=1, num_dots=2)
    with pytest.raises(RuntimeError):
        port.recon_downsampled(fr, 2)
