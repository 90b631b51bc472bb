"""Host entropy decoding of AC groups (pikgpu_decode_ac_pass) on the CPU:
bit-exact against the reference decoder on the committed bitstreams
(tests/golden/ac_groups.npz, written and decoded by the reference), live
against the reference encoder + decoder on more frames, and on corrupt
streams (the reference decoder and this one must both reject a stream or
both decode it to the same coefficients)."""
import ctypes
import json
import os
import tempfile

import numpy as np
import pytest

from pik_amd import abi, bitstream, synth
from pik_amd.runtime import PikGpuError

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "ac_groups.npz")


def _cases():
    z = np.load(FIXTURE)
    return z, sorted({k.split("/")[0] for k in z.files})


def _groups(z, name):
    data = z[name + "/data"].tobytes()
    out, pos = [], 0
    for s in z[name + "/sizes"]:
        out.append(data[pos:pos + int(s)])
        pos += int(s)
    return out


def _recipe(z, name):
    return json.loads(str(z[name + "/recipe"]))


@pytest.mark.parametrize("name", _cases()[1])
@pytest.mark.parametrize("threads", [1, 3])
def test_fixture_bit_exact(name, threads):
    z, _ = _cases()
    rec = _recipe(z, name)
    d = bitstream.decode_ac_pass(rec["w"], rec["h"], z[name + "/acs"],
                                 _groups(z, name), threads=threads)
    assert np.array_equal(d["ac"], z[name + "/ac"])
    assert np.array_equal(d["noise"].view(np.uint32),
                          z[name + "/noise"].view(np.uint32))
    assert np.array_equal(d["consumed"], z[name + "/consumed"])
    # every byte of each group is read
    assert np.array_equal(d["consumed"], z[name + "/sizes"])


@pytest.mark.parametrize("frac", [0.5, 0.9])
def test_truncated_streams_are_rejected(frac):
    z, names = _cases()
    for name in names:
        if name == "zero_ac":  # 8 bytes, the last one padding
            continue
        rec = _recipe(z, name)
        groups = [g[:int(len(g) * frac)] for g in _groups(z, name)]
        with pytest.raises(PikGpuError) as e:
            bitstream.decode_ac_pass(rec["w"], rec["h"], z[name + "/acs"],
                                     groups)
        assert e.value.code == abi.EDATA
        assert "group" in str(e.value)


def test_error_names_lowest_bad_group():
    z, _ = _cases()
    rec = _recipe(z, "groups_fast")
    groups = _groups(z, "groups_fast")
    groups[4] = groups[4][:10]
    groups[2] = groups[2][:5]
    with pytest.raises(PikGpuError) as e:
        bitstream.decode_ac_pass(rec["w"], rec["h"], z["groups_fast/acs"],
                                 groups, threads=4)
    assert "group 2:" in str(e.value)


def test_argument_checks():
    z, _ = _cases()
    rec = _recipe(z, "lines_groups")
    groups = _groups(z, "lines_groups")
    acs = z["lines_groups/acs"]
    lib = abi.load()
    by, bx = acs.shape
    out = np.zeros((3, by, bx * 64), np.int16)
    bufs = [np.frombuffer(g, np.uint8) for g in groups]
    ptrs = np.array([b.ctypes.data for b in bufs], np.uintp)
    sizes = np.array([len(g) for g in groups], np.uintp)
    acs_c = np.ascontiguousarray(acs)

    def call(**over):
        p = abi.AcPassC()
        p.xsize, p.ysize = rec["w"], rec["h"]
        p.ac_strategy = acs_c.ctypes.data
        p.ac_strategy_stride = bx
        p.num_groups = len(groups)
        p.group_data = ptrs.ctypes.data
        p.group_size = sizes.ctypes.data
        for c in range(3):
            p.ac[c] = out[c].ctypes.data
        p.ac_stride = bx * 64
        for k, v in over.items():
            setattr(p, k, v)
        msg = ctypes.create_string_buffer(128)
        return lib.pikgpu_decode_ac_pass(ctypes.byref(p), 1, msg, 128), \
            msg.value.decode()

    assert call()[0] == abi.OK
    assert np.array_equal(out, z["lines_groups/ac"])
    assert call(num_groups=3)[0] == abi.EINVAL
    assert call(ac_stride=bx * 64 - 1)[0] == abi.EINVAL
    assert call(ac_strategy_stride=bx - 1)[0] == abi.EINVAL
    assert call(xsize=0)[0] == abi.EINVAL
    assert call(group_data=None)[0] == abi.EINVAL
    assert lib.pikgpu_decode_ac_pass(None, 1, None, 0) == abi.EINVAL


def test_inconsistent_strategy_is_rejected():
    z, _ = _cases()
    rec = _recipe(z, "small_noise")
    acs = z["small_noise/acs"].copy()
    # a 16x16 transform's second block claiming to start a new one
    ys, xs = np.nonzero(acs == (abi.ACS_DCT16X16 << 4) | 1)
    assert len(ys)
    acs[ys[0], xs[0]] = abi.ACS_DCT16X16 << 4
    with pytest.raises(PikGpuError) as e:
        bitstream.decode_ac_pass(rec["w"], rec["h"], acs,
                                 _groups(z, "small_noise"))
    assert e.value.code == abi.EDATA
    assert "strategy" in str(e.value)


def _need_ref():
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    return ref


@pytest.mark.parametrize("w,h,seed,kw,enc", [
    (72, 520, 12, {}, {}), (520, 72, 13, dict(lines=0.3), dict(fast_mode=True)),
    (1100, 600, 14, dict(distance=0.5), {}),
    (333, 333, 15, dict(distance=3.0, grayscale=True), {}),
    (400, 300, 16, dict(line_art=True), dict(noise=(-0.5, 1.0, 0.25)))])
def test_live_against_reference(w, h, seed, kw, enc):
    ref = _need_ref()
    fr = synth.make_frame(w, h, seed=seed, **kw)
    groups = ref.encode_ac_groups(fr, noise=enc.get("noise"),
                                  fast_mode=enc.get("fast_mode", False))
    want, noise, used = ref.decode_ac_groups(fr, groups, check_dequant=True)
    d = bitstream.decode_ac_pass(w, h, fr.ac_strategy, groups)
    assert np.array_equal(d["ac"], want)
    assert np.array_equal(d["noise"].view(np.uint32), noise.view(np.uint32))
    assert np.array_equal(d["consumed"], used)


def _ref_decode_isolated(ref, frame, groups, tmp):
    """The reference decoder in a child process: on some corrupt streams it
    stops the process with an assertion (pik/entropy_coder.h:144)."""
    path = os.path.join(tmp, "r.npz")
    pid = os.fork()
    if pid == 0:
        code = 1
        try:
            ac, _, used = ref.decode_ac_groups(frame, groups,
                                               check_dequant=False)
            np.savez(path, ac=ac, used=used)
            code = 0
        except RuntimeError:
            pass
        os._exit(code)
    _, st = os.waitpid(pid, 0)
    if os.WIFSIGNALED(st):
        return "abort", None
    if os.WEXITSTATUS(st) != 0:
        return "fail", None
    q = np.load(path)
    return "ok", (q["ac"], q["used"])


@pytest.mark.parametrize("name", ["small_noise", "small_fast", "gray",
                                  "dense_ac"])
def test_corrupt_streams_match_reference(name):
    ref = _need_ref()
    z, _ = _cases()
    rec = _recipe(z, name)
    fr = synth.make_frame(rec["w"], rec["h"], seed=rec["seed"], **rec["kw"])
    fr.ac_strategy = z[name + "/acs"]
    data = _groups(z, name)[0]
    rng = np.random.default_rng(7)
    seen = set()
    with tempfile.TemporaryDirectory() as tmp:
        for trial in range(45):
            b = bytearray(data)
            mode = trial % 3
            if mode == 0:
                b[int(rng.integers(len(b)))] ^= 1 << int(rng.integers(8))
            elif mode == 1:
                b[int(rng.integers(len(b)))] = int(rng.integers(256))
            else:
                b = b[:int(rng.integers(len(b)))]
            groups = [bytes(b)]
            rs, rv = _ref_decode_isolated(ref, fr, groups, tmp)
            try:
                d = bitstream.decode_ac_pass(fr.xsize, fr.ysize,
                                             fr.ac_strategy, groups)
                ms = "ok"
            except PikGpuError as e:
                assert e.code == abi.EDATA
                ms = "fail"
            seen.add((rs, ms))
            if rs == "abort":  # the reference has no defined result
                continue
            assert rs == ms, (trial, rs, ms)
            if rs == "ok":
                assert np.array_equal(rv[0], d["ac"]), trial
                assert np.array_equal(rv[1], d["consumed"]), trial
    assert ("fail", "fail") in seen
