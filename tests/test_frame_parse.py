"""Host parsing of whole frames (pikgpu_parse_frame) on the CPU: bit-exact
against the reference decoder on the committed frames (tests/golden/
frames.npz), live against the reference's encoder and decoder pieces on more
frames, on frame headers written here (extensions with bits to skip, the
refused and invalid encodings), and on corrupt frames (both decoders reject
a frame or both give the same result)."""
import ctypes
import json
import os
import tempfile

import numpy as np
import pytest

from pik_amd import abi, bitstream, synth
from pik_amd.runtime import PikGpuError

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "frames.npz")
FIELDS = ("dc", "ac_strategy", "quant_field", "sigma_lut_ids", "ytox_map",
          "ytob_map", "dequant_cf", "dequant_map")
HEADER = ("flags", "gaborish", "predict_lf", "predict_hf",
          "adaptive_reconstruction", "epf_enable_adaptive", "epf_sigma",
          "epf_use_sharpened", "num_passes", "global_scale", "quant_dc",
          "ytox_dc", "ytob_dc", "num_table_sets", "num_dots", "frame_size")


def _cases():
    z = np.load(FIXTURE)
    return z, sorted({k.split("/")[0] for k in z.files})


def _header_of(frame, info):
    """The parsed frame's header fields in HEADER order."""
    sets = frame.dequant_offsets.size // 21
    return [frame.flags, frame.gaborish, frame.predict_lf, frame.predict_hf,
            frame.adaptive_reconstruction, frame.epf_enable_adaptive,
            info["epf_sigma"], frame.epf_use_sharpened,
            len(info["passes"]), frame.global_scale, frame.quant_dc,
            frame.ytox_dc, frame.ytob_dc, sets, len(frame.dots),
            info["frame_size"] & 0x7FFFFFFF]


def _compare(frame, info, want):
    """want: the reference decoder's result (fixture arrays or
    oracle.ref.decode_frame's dict)."""
    for k in FIELDS:
        assert np.array_equal(getattr(frame, k), want[k]), k
    assert [int(v) for v in _header_of(frame, info)] == \
        [int(want[k]) for k in HEADER]
    assert np.array_equal(frame.dots, want["dots"])
    assert len(info["passes"]) == len(want["passes"])
    for a, b in zip(info["passes"], want["passes"]):
        assert np.array_equal(a, b)
    assert np.array_equal(info["noise"], want["noise"])
    assert info["consumed"] == int(want["consumed"])


def _fixture_want(z, name):
    w = {k: z[name + "/" + k] for k in FIELDS + ("dots", "noise")}
    w.update(zip(HEADER, z[name + "/header"]))
    w["passes"] = list(z[name + "/passes"])
    w["consumed"] = int(z[name + "/consumed"])
    return w


@pytest.mark.parametrize("name", _cases()[1])
@pytest.mark.parametrize("threads", [1, 4])
def test_fixture_bit_exact(name, threads):
    z, _ = _cases()
    r = json.loads(str(z[name + "/recipe"]))
    data = z[name + "/data"].tobytes()
    frame, info = bitstream.parse_frame(data, r["w"], r["h"], threads=threads)
    _compare(frame, info, _fixture_want(z, name))
    # the default tables, once per encoded table set
    sets = int(z[name + "/header"][HEADER.index("num_table_sets")])
    assert frame.dequant_offsets.size == 21 * sets


def test_fixture_cases_cover_the_format():
    z, names = _cases()
    hdr = {n: dict(zip(HEADER, z[n + "/header"])) for n in names}
    assert any(h["num_passes"] > 1 for h in hdr.values())
    assert any(h["num_table_sets"] > 1 for h in hdr.values())
    assert any(h["num_dots"] > 0 for h in hdr.values())
    assert any(h["flags"] & abi.FLAG_GRAYSCALE_OPT for h in hdr.values())
    assert any(not h["adaptive_reconstruction"] for h in hdr.values())
    assert any(h["epf_use_sharpened"] for h in hdr.values())
    dots = np.concatenate([z[n + "/dots"] for n in names])
    assert dots["transform"].any() and not dots["transform"].all()
    assert any(np.abs(z[n + "/noise"]).max() > 0 for n in names)
    assert any(z[n + "/dequant_map"].max() > 0 for n in names)


def test_trailing_bytes_are_not_read():
    z, _ = _cases()
    r = json.loads(str(z["dots/recipe"]))
    data = z["dots/data"].tobytes()
    frame, info = bitstream.parse_frame(data + b"\x55" * 100, r["w"], r["h"])
    assert info["consumed"] == len(data)
    _compare(frame, info, _fixture_want(z, "dots"))


def test_truncated_frames_are_rejected():
    z, _ = _cases()
    r = json.loads(str(z["passes/recipe"]))
    data = z["passes/data"].tobytes()
    for cut in (len(data) - 1, len(data) // 2, 40, 3):
        with pytest.raises(PikGpuError) as e:
            bitstream.parse_frame(data[:cut], r["w"], r["h"])
        assert e.value.code == abi.EDATA


def test_argument_checks():
    lib = abi.load()
    h = ctypes.c_void_p()
    msg = ctypes.create_string_buffer(64)
    buf = (ctypes.c_uint8 * 4)()
    assert lib.pikgpu_parse_frame(buf, 4, 8, 8, 1, None, msg, 64) == abi.EINVAL
    assert lib.pikgpu_parse_frame(None, 4, 8, 8, 1, ctypes.byref(h), msg,
                                  64) == abi.EINVAL
    assert lib.pikgpu_parse_frame(buf, 4, 0, 8, 1, ctypes.byref(h), msg,
                                  64) == abi.EINVAL
    assert lib.pikgpu_parsed_get(None, None, None) == abi.EINVAL
    lib.pikgpu_parsed_free(None)


# ---------------------------------------------------------------------------
# Frame headers written here. The writer picks, per U32 field, the shortest
# encoding of the distribution (the lowest selector on ties), as the
# reference's ChooseEncoding does; test_header_writer_matches_reference pins
# it against the reference encoder's header bytes.

class BitWriter:
    def __init__(self):
        self.bits = []

    def write(self, n, v):
        self.bits += [(v >> i) & 1 for i in range(n)]

    def u32(self, dist, v):
        if dist > 0xFFFFFFDF:  # raw bits
            self.write(dist - 0xFFFFFFDF, v)
            return
        best = None
        for sel in range(4):
            b = (dist >> (8 * sel)) & 0xFF
            if b & 0x80:
                if v == b & 0x7F:
                    cand = (0, sel, 0, 0)
                else:
                    continue
            elif b & 0x40:
                off, nb = ((b >> 3) & 7) + 1, (b & 7) + 1
                if not off <= v < off + (1 << nb):
                    continue
                cand = (nb, sel, nb, v - off)
            else:
                if v >= 1 << b:
                    continue
                cand = (b, sel, b, v)
            if best is None or cand[0] < best[0]:
                best = cand
        assert best is not None, (hex(dist), v)
        self.write(2, best[1])
        self.write(best[2], best[3])

    def u64(self, v):
        if v == 0:
            self.write(2, 0)
        elif v <= 16:
            self.write(2, 1)
            self.write(4, v - 1)
        elif v <= 272:
            self.write(2, 2)
            self.write(8, v - 17)
        else:
            self.write(2, 3)
            self.write(12, v & 4095)
            v >>= 12
            shift = 12
            while v:
                self.write(1, 1)
                if shift == 60:
                    self.write(4, v & 15)
                    break
                self.write(8, v & 255)
                v >>= 8
                shift += 8
            else:
                self.write(1, 0)

    def boolean(self, b):
        self.write(1, int(bool(b)))

    def to_bytes(self):
        bits = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(sum(bits[i + j] << j for j in range(8))
                     for i in range(0, len(bits), 8))


def frame_header(encoding=0, has_alpha=0, flags=0, gaborish=2, predict=(1, 1),
                 ar=1, epf_adaptive=1, sigma=0, sharpened=0, passes=1,
                 downsampling=(), ext_bits=None, frame_info=None):
    """FrameHeader bytes (pik/headers.h:201-251), byte aligned. ext_bits:
    None (no extensions) or the number of extension bits to append (ones
    and zeros, which every decoder skips)."""
    w = BitWriter()
    w.u64(0)
    w.boolean(has_alpha)
    if frame_info is None:
        w.boolean(1)
    else:
        w.boolean(0)
        w.u32(0x20088180, frame_info[0])
        w.boolean(1)
        w.u32(0xFFFFFFDF + 32, frame_info[1])
        w.boolean(1)
    w.u32(0x51828180, encoding)
    if encoding == 0:
        w.u32(0x20181008, flags)
        w.u32(0x51828180, gaborish)
        w.boolean(predict[0])
        w.boolean(predict[1])
        w.boolean(ar)
        if ar:
            w.boolean(epf_adaptive)
            if not epf_adaptive:
                w.u32(0x0A090880, sigma)
            w.boolean(sharpened)
        w.u32(0x20181008, passes)
        w.u32(0x83828180, len(downsampling))
        for f, p in downsampling:
            w.u32(0x88848281, f)
            w.u32(0x52828180, p)
    if encoding == 2:
        w.boolean(0)
        w.boolean(0)
    if ext_bits is None:
        w.u64(0)
    else:
        w.u64(5)
        w.u64(ext_bits)
        for i in range(ext_bits):
            w.write(1, (i * 7 + 3) % 3 == 0)
    return w.to_bytes()


def _tiny():
    """The 'tiny' fixture frame and its body after the header."""
    z, _ = _cases()
    data = z["tiny/data"].tobytes()
    hdr = frame_header()
    assert data.startswith(hdr)
    return z, data, data[len(hdr):]


def _need_ref():
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    return ref


def test_header_writer_matches_reference():
    # The fixture frames' headers are the reference's WritePassHeader bytes.
    z, data, body = _tiny()
    assert len(body) < len(data)
    d = z["dots/data"].tobytes()
    assert d.startswith(frame_header(frame_info=(7, 0x01020304)))


@pytest.mark.parametrize("ext_bits", [0, 13, 40, 300])
def test_extension_bits_are_skipped(ext_bits):
    z, data, body = _tiny()
    frame, info = bitstream.parse_frame(frame_header(ext_bits=ext_bits) + body,
                                        13, 7)
    want = _fixture_want(z, "tiny")
    want["consumed"] = int(want["consumed"]) + len(
        frame_header(ext_bits=ext_bits)) - len(frame_header())
    _compare(frame, info, want)


def test_extensions_against_reference():
    ref = _need_ref()
    _, _, body = _tiny()
    for ext_bits in (0, 13, 300):
        data = frame_header(ext_bits=ext_bits,
                            downsampling=((2, 0), (8, 9))) + body
        want = ref.decode_frame(data, 13, 7)
        frame, info = bitstream.parse_frame(data, 13, 7)
        _compare(frame, info, want)
        assert info["downsampling"] == [(2, 0), (8, 9)]


@pytest.mark.parametrize("kw,code", [
    (dict(encoding=2), abi.EUNSUPPORTED),       # lossless
    (dict(has_alpha=1), abi.EUNSUPPORTED),
    (dict(flags=abi.FLAG_GRADIENT_MAP), abi.EUNSUPPORTED),
    (dict(passes=17), abi.EUNSUPPORTED),
    (dict(encoding=1), abi.EDATA),              # progressive
    (dict(encoding=5), abi.EDATA),
    (dict(passes=0), abi.EDATA),
    (dict(gaborish=5), abi.EDATA),
])
def test_refused_headers(kw, code):
    _, _, body = _tiny()
    with pytest.raises(PikGpuError) as e:
        bitstream.parse_frame(frame_header(**kw) + body, 13, 7)
    assert e.value.code == code, str(e.value)


def test_header_flags_that_decode_but_do_not_reconstruct():
    # Noise and fixed-sigma EPF parse (the reconstruction refuses them).
    _, _, body = _tiny()
    frame, _ = bitstream.parse_frame(
        frame_header(flags=abi.FLAG_NOISE) + body, 13, 7)
    assert frame.flags == abi.FLAG_NOISE
    frame, info = bitstream.parse_frame(
        frame_header(epf_adaptive=0, sigma=70) + body, 13, 7)
    assert frame.epf_enable_adaptive == 0 and info["epf_sigma"] == 70


def test_custom_tables_are_refused():
    _, data, body = _tiny()
    hdr = frame_header()
    assert body[0] == 0  # no encoded tables
    for mode in range(1, 7):
        with pytest.raises(PikGpuError) as e:
            bitstream.parse_frame(hdr + bytes([1, mode]) + body[1:], 13, 7)
        assert e.value.code == abi.EUNSUPPORTED
    for bad in (bytes([1, 0, 1]), bytes([2, 0, 0, 7, 1]), bytes([1, 9])):
        # a second predefined table, a copy of itself, an unknown mode
        with pytest.raises(PikGpuError) as e:
            bitstream.parse_frame(hdr + bad + body[1:], 13, 7)
        assert e.value.code == abi.EDATA


# ---------------------------------------------------------------------------
# Live against the reference.

LIVE = [
    (64, 64, 1, dict(num_dots=3, lut_ids="rand"), 1, 0),
    (520, 600, 5, dict(lines=0.3), 0, 2),
    (1100, 300, 6, dict(distance=2.5, num_dots=8), 3, 1),
    (9, 1030, 7, dict(grayscale=True), 0, 0),
]


@pytest.mark.parametrize("w,h,seed,kw,passes,variant", LIVE)
def test_live_against_reference(w, h, seed, kw, passes, variant):
    ref = _need_ref()
    fr = synth.make_frame(w, h, seed=seed, **kw)
    if fr.grayscale:
        fr.flags |= abi.FLAG_GRAYSCALE_OPT
    extra = synth.make_extra_passes(fr, passes, seed=seed)
    bits = [ref.VAR_STATIC_DOTS, ref.VAR_GROUP_HEADERS | ref.VAR_FAST_AC,
            ref.VAR_COPY_TABLE | ref.VAR_EXTENSIONS]
    data = ref.encode_frame(fr, extra, None, bits[variant % 3])
    want = ref.decode_frame(data, w, h)
    frame, info = bitstream.parse_frame(data, w, h)
    _compare(frame, info, want)
    assert info["consumed"] == len(data)


def _ref_isolated(ref, w, h, data, tmp):
    path = os.path.join(tmp, "r.npz")
    pid = os.fork()
    if pid == 0:
        code = 1
        try:
            r = ref.decode_frame(data, w, h)
            arrays = {k: v for k, v in r.items()
                      if isinstance(v, np.ndarray)}
            arrays["passes"] = np.stack(r["passes"])
            np.savez(path, header=np.array([r[k] for k in HEADER]),
                     consumed=r["consumed"], **arrays)
            code = 0
        except RuntimeError:
            pass
        os._exit(code)
    _, st = os.waitpid(pid, 0)
    if os.WIFSIGNALED(st):
        return "abort", None
    if os.WEXITSTATUS(st) != 0:
        return "fail", None
    z = np.load(path)
    want = {k: z[k] for k in z.files if k not in ("header", "passes")}
    want.update(zip(HEADER, z["header"]))
    want["passes"] = list(z["passes"])
    return "ok", want


@pytest.mark.parametrize("name", ["dots", "passes", "tables"])
def test_corrupt_frames_match_reference(name):
    ref = _need_ref()
    z, _ = _cases()
    r = json.loads(str(z[name + "/recipe"]))
    data = z[name + "/data"].tobytes()
    rng = np.random.default_rng(5)
    seen = set()
    with tempfile.TemporaryDirectory() as tmp:
        for trial in range(45):
            b = bytearray(data)
            mode = trial % 3
            # bias towards the pass-global sections ahead of the groups
            span = len(b) if trial % 2 else min(len(b), 600)
            if mode == 0:
                b[int(rng.integers(span))] ^= 1 << int(rng.integers(8))
            elif mode == 1:
                b[int(rng.integers(span))] = int(rng.integers(256))
            else:
                b = b[:int(rng.integers(len(b)))]
            rs, rv = _ref_isolated(ref, r["w"], r["h"], bytes(b), tmp)
            try:
                frame, info = bitstream.parse_frame(bytes(b), r["w"], r["h"])
                ms = "ok"
            except PikGpuError as e:
                if e.code == abi.EUNSUPPORTED:
                    continue
                assert e.code == abi.EDATA
                ms = "fail"
            seen.add((rs, ms))
            if rs == "abort":
                continue
            assert rs == ms, (trial, rs, ms)
            if rs == "ok":
                _compare(frame, info, rv)
    assert ("fail", "fail") in seen


def test_reused_parse_storage():
    # A parsed frame handed back in is refilled: parsing a different frame
    # into it gives what a fresh parse gives, and a failed parse leaves the
    # handle usable.
    z, _ = _cases()
    lib = abi.load()
    h = ctypes.c_void_p()
    msg = ctypes.create_string_buffer(128)
    for name in ("passes", "dots", "tiny", "plain", "dots"):
        r = json.loads(str(z[name + "/recipe"]))
        raw = np.frombuffer(z[name + "/data"].tobytes(), np.uint8)
        assert lib.pikgpu_parse_frame(raw.ctypes.data, raw.size, r["w"],
                                      r["h"], 2, ctypes.byref(h), msg,
                                      128) == abi.OK
        fc = abi.FrameC()
        info = abi.ParsedInfoC()
        lib.pikgpu_parsed_get(h, ctypes.byref(fc), ctypes.byref(info))
        assert info.consumed == raw.size and fc.num_dots == len(
            z[name + "/dots"])
        by, bx = -(-r["h"] // 8), -(-r["w"] // 8)
        for c in range(3):
            ac = np.ctypeslib.as_array(
                ctypes.cast(fc.ac[c], ctypes.POINTER(ctypes.c_int16)),
                (by * bx * 64,))
            assert np.array_equal(ac.reshape(by, bx * 64),
                                  z[name + "/passes"][0][c])
        bad = raw[:raw.size // 2].copy()
        assert lib.pikgpu_parse_frame(bad.ctypes.data, bad.size, r["w"],
                                      r["h"], 2, ctypes.byref(h), msg,
                                      128) == abi.EDATA
        assert h.value
    lib.pikgpu_parsed_free(h)
