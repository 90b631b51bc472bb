"""Host decoding of DC groups sections (pikgpu_decode_dc_groups) on the
CPU: bit-exact against the reference decoder on the committed sections
(tests/golden/dc_groups.npz), live against the reference's encoder and
decoder pieces on more frames, and on corrupt sections (both decoders
reject a section or both give the same result)."""
import ctypes
import json
import os
import tempfile

import numpy as np
import pytest

from pik_amd import abi, bitstream, synth
from pik_amd.runtime import PikGpuError

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "dc_groups.npz")
FIELDS = ("dc", "ac_strategy", "quant_field", "sigma_lut_ids")


def _cases():
    z = np.load(FIXTURE)
    return z, sorted({k.split("/")[0] for k in z.files})


def _recipe(z, name):
    return json.loads(str(z[name + "/recipe"]))


@pytest.mark.parametrize("name", _cases()[1])
@pytest.mark.parametrize("threads", [1, 4])
def test_fixture_bit_exact(name, threads):
    z, _ = _cases()
    r = _recipe(z, name)
    got = bitstream.decode_dc_groups(r["w"], r["h"], z[name + "/data"].tobytes(),
                                     threads=threads)
    for k in FIELDS:
        assert np.array_equal(got[k], z[name + "/" + k]), k
    assert got["consumed"] == int(z[name + "/consumed"])


def test_predictor_tie_and_saturation_cases_are_the_default_builds():
    # DC steps up to 10000: the predictor costs reach 32768, where the
    # reference's default (AVX2) build and its portable build choose
    # differently; the fixture holds the default build's stream.
    z, _ = _cases()
    r = _recipe(z, "wide_dc")
    assert r["opts"]["kind"] == "avx2"
    got = bitstream.decode_dc_groups(r["w"], r["h"], z["wide_dc/data"].tobytes())
    assert np.abs(got["dc"].astype(np.int32)).max() > 9000
    assert np.array_equal(got["dc"], z["wide_dc/dc"])


@pytest.mark.parametrize("name", ["small_noise", "two_groups", "groups_fast"])
def test_truncated_sections_are_rejected(name):
    z, _ = _cases()
    r = _recipe(z, name)
    data = z[name + "/data"].tobytes()
    for cut in (len(data) // 2, len(data) - 1):
        with pytest.raises(PikGpuError) as e:
            bitstream.decode_dc_groups(r["w"], r["h"], data[:cut])
        assert e.value.code == abi.EDATA


def test_argument_checks():
    lib = abi.load()
    z, _ = _cases()
    data = z["tiny/data"].tobytes()
    raw = np.frombuffer(data, np.uint8)
    dc = np.zeros((3, 1, 2), np.int16)
    acs = np.zeros((1, 2), np.uint8)
    qf = np.zeros((1, 2), np.int32)
    lut = np.zeros((1, 2), np.uint8)

    def call(**over):
        s = abi.DcSectionC()
        s.xsize, s.ysize = 13, 7
        s.data, s.size = raw.ctypes.data, len(data)
        for c in range(3):
            s.dc[c] = dc[c].ctypes.data
        s.dc_stride = 2
        s.ac_strategy, s.ac_strategy_stride = acs.ctypes.data, 2
        s.quant_field, s.quant_field_stride = qf.ctypes.data, 2
        s.sigma_lut_ids, s.sigma_lut_stride = lut.ctypes.data, 2
        for k, v in over.items():
            setattr(s, k, v)
        msg = ctypes.create_string_buffer(128)
        return lib.pikgpu_decode_dc_groups(ctypes.byref(s), 1, msg, 128), s

    rc, s = call()
    assert rc == abi.OK and s.consumed == len(data)
    assert np.array_equal(dc, z["tiny/dc"])
    assert call(dc_stride=1)[0] == abi.EINVAL
    assert call(quant_field=None)[0] == abi.EINVAL
    assert call(ysize=0)[0] == abi.EINVAL
    assert lib.pikgpu_decode_dc_groups(None, 1, None, 0) == abi.EINVAL


def _need_ref():
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    return ref


@pytest.mark.parametrize("w,h,seed,kw", [
    (2056, 2060, 20, dict(distance=3.0)), (9, 2100, 21, {}),
    (517, 333, 22, dict(line_art=True, lut_ids="rand")),
    (64, 64, 23, dict(distance=0.5, grayscale=True))])
def test_live_against_reference(w, h, seed, kw):
    ref = _need_ref()
    fr = synth.make_frame(w, h, seed=seed, **kw)
    sec = ref.encode_dc_groups(fr)
    want = ref.decode_dc_groups(w, h, sec)
    got = bitstream.decode_dc_groups(w, h, sec)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k
    assert got["consumed"] == want["consumed"] == len(sec)


def _ref_isolated(ref, w, h, data, tmp):
    path = os.path.join(tmp, "r.npz")
    pid = os.fork()
    if pid == 0:
        code = 1
        try:
            r = ref.decode_dc_groups(w, h, data)
            np.savez(path, consumed=r.pop("consumed"), **r)
            code = 0
        except RuntimeError:
            pass
        os._exit(code)
    _, st = os.waitpid(pid, 0)
    if os.WIFSIGNALED(st):
        return "abort", None
    if os.WEXITSTATUS(st) != 0:
        return "fail", None
    return "ok", dict(np.load(path))


@pytest.mark.parametrize("name", ["small_noise", "lines_lut", "narrow"])
def test_corrupt_sections_match_reference(name):
    ref = _need_ref()
    z, _ = _cases()
    r = _recipe(z, name)
    data = z[name + "/data"].tobytes()
    rng = np.random.default_rng(11)
    seen = set()
    with tempfile.TemporaryDirectory() as tmp:
        for trial in range(40):
            b = bytearray(data)
            mode = trial % 3
            if mode == 0:
                b[int(rng.integers(len(b)))] ^= 1 << int(rng.integers(8))
            elif mode == 1:
                b[int(rng.integers(len(b)))] = int(rng.integers(256))
            else:
                b = b[:int(rng.integers(len(b)))]
            rs, rv = _ref_isolated(ref, r["w"], r["h"], bytes(b), tmp)
            try:
                got = bitstream.decode_dc_groups(r["w"], r["h"], bytes(b))
                ms = "ok"
            except PikGpuError as e:
                assert e.code == abi.EDATA
                ms = "fail"
            seen.add((rs, ms))
            if rs == "abort":
                continue
            assert rs == ms, (trial, rs, ms)
            if rs == "ok":
                for k in FIELDS:
                    assert np.array_equal(got[k], rv[k]), (trial, k)
                assert got["consumed"] == int(rv["consumed"])
    assert ("fail", "fail") in seen
