"""pikgpu_reconstruct_noalloc: the reconstruction without any device
allocation (every buffer in the library's fixed device scratch array)
against pikgpu_reconstruct, bit for bit, on two calls in a row, the second
with a smaller frame; frames too large for the scratch are refused before
the device is touched."""
import ctypes

import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import Reconstructor

pytestmark = pytest.mark.gpu


def noalloc(frame, stream=None):
    """Calls the entry point; returns (rc, linear (3, H, W), message)."""
    lib = abi.load()
    fc, keep = frame.to_c()
    out = np.full((3, frame.ysize, frame.xsize), np.nan, np.float32)
    planes = (abi.c_float_p * 3)(
        *[out[c].ctypes.data_as(abi.c_float_p) for c in range(3)])
    msg = ctypes.create_string_buffer(256)
    rc = lib.pikgpu_reconstruct_noalloc(ctypes.byref(fc), planes,
                                        frame.xsize * 4, stream, msg,
                                        len(msg))
    return rc, out, msg.value.decode()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def blank_frame(w, h):
    """A decodable frame of any size without synthesising an image: zero
    coefficients, DCT8 everywhere, the library tables."""
    t = synth.make_frame(8, 8, seed=1)
    bx, by = -(-w // 8), -(-h // 8)
    tx, ty = -(-w // 64), -(-h // 64)
    f = t.copy()
    f.xsize, f.ysize = w, h
    f.ac = np.zeros((3, by, bx * 64), np.int16)
    f.dc = np.zeros((3, by, bx), np.int16)
    f.quant_field = np.ones((by, bx), np.int32)
    f.ac_strategy = np.zeros((by, bx), np.uint8)
    f.sigma_lut_ids = np.zeros((by, bx), np.uint8)
    f.dequant_cf = np.zeros((ty, tx), np.uint8)
    f.ytox_map = np.full((ty, tx), 128, np.int32)
    f.ytob_map = np.full((ty, tx), 120, np.int32)
    f.dots = f.dots[:0]
    return f.contiguous()


def test_matches_reconstruct_on_consecutive_calls():
    first = synth.make_frame(1920, 1080, seed=201, num_dots=6,
                             lut_ids="rand", distance=1.5)
    second = synth.make_frame(333, 257, seed=202, num_dots=2)
    print("no-alloc: 1920x1080 then 333x257", flush=True)
    with Reconstructor(0) as R:
        want = [R.reconstruct(first), R.reconstruct(second)]
        stream = R.stream
        rc, got, msg = noalloc(first)
        assert rc == abi.OK, msg
        assert np.array_equal(bits(got), bits(want[0]))
        # the second call, smaller, on a caller's stream: nothing of the
        # first frame may leak into it
        rc, got, msg = noalloc(second, stream)
        assert rc == abi.OK, msg
        assert np.array_equal(bits(got), bits(want[1]))
        # and the first again after the smaller one
        rc, got, msg = noalloc(first)
        assert rc == abi.OK, msg
        assert np.array_equal(bits(got), bits(want[0]))


def test_refuses_frames_beyond_the_scratch():
    print("no-alloc: size limit", flush=True)
    rc, _, msg = noalloc(blank_frame(2048, 2048))
    assert rc == abi.EINVAL
    assert "scratch" in msg
    bad = synth.make_frame(64, 64, seed=3)
    bad.quant_field[0, 0] = 0
    rc, _, msg = noalloc(bad)
    assert rc == abi.EINVAL and "quant field" in msg
