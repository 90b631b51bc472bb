"""ctypes bindings to the reference harness libraries under oracle/_ref
(built from the google/pik sources by oracle/Makefile together with
ref_harness.cc). Test infrastructure: the reference is the checker, or the
CPU baseline timed beside the GPU, never part of the product."""
import ctypes
import os

import numpy as np

from pik_amd import abi
from pik_amd.frame import div_ceil

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
SCALAR = os.path.join(REF_DIR, "libpikref_scalar.so")
AVX2 = os.path.join(REF_DIR, "libpikref_avx2.so")

_libs = {}


def available(kind="scalar"):
    return os.path.exists(SCALAR if kind == "scalar" else AVX2)


def lib(kind="scalar"):
    if kind in _libs:
        return _libs[kind]
    path = SCALAR if kind == "scalar" else AVX2
    if not os.path.exists(path):
        raise RuntimeError("reference harness not built: %s" % path)
    L = ctypes.CDLL(path)
    L.pikref_last_error.restype = ctypes.c_char_p
    L.pikref_recon.argtypes = [ctypes.POINTER(abi.FrameC), ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_void_p]
    L.pikref_default_tables.argtypes = [
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.pikref_recon_passes.argtypes = [ctypes.POINTER(abi.FrameC), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_size_t, ctypes.c_void_p,
                                      ctypes.c_void_p]
    L.pikref_recon_downsampled.argtypes = [ctypes.POINTER(abi.FrameC),
                                           ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p]
    L.pikref_epf_weight.restype = ctypes.c_float
    L.pikref_epf_weight.argtypes = [ctypes.c_int, ctypes.c_int]
    L.pikref_transform.argtypes = [ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p]
    L.pikref_gaborish.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.c_size_t,
                                  ctypes.c_void_p]
    L.pikref_opsin_to_linear.argtypes = [ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_size_t]
    L.pikref_linear_to_xyb.argtypes = [ctypes.c_float] * 3 + [ctypes.c_void_p]
    L.pikref_quantize.argtypes = [ctypes.POINTER(abi.FrameC), ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
    _libs[kind] = L
    return L


def default_tables(kind="scalar"):
    L = lib(kind)
    ts, no = ctypes.c_size_t(), ctypes.c_size_t()
    L.pikref_default_tables(None, 0, None, 0, ctypes.byref(ts), ctypes.byref(no))
    t = np.zeros(ts.value, np.float32)
    o = np.zeros(no.value, np.uint32)
    L.pikref_default_tables(t.ctypes.data, t.size, o.ctypes.data, o.size,
                            ctypes.byref(ts), ctypes.byref(no))
    return t, o


def stage_shape(frame, stage):
    bx, by = frame.xsize_blocks, frame.ysize_blocks
    if stage in (abi.STAGE_DC_DEQUANT, abi.STAGE_DC):
        return (by, bx)
    if stage == abi.STAGE_COEFFS:
        return (by, bx * 64)
    if stage == abi.STAGE_LINEAR:
        return (frame.ysize, frame.xsize)
    return (by * 8, bx * 8)


def recon(frame, threads=0, stages=True, kind="scalar"):
    """Runs the reference reconstruction. Returns (linear (3,H,W) float32,
    dict stage_name -> (3, h, w) float32)."""
    L = lib(kind)
    fc, keep = frame.to_c()
    out = np.zeros((3, frame.ysize, frame.xsize), np.float32)
    out_ptrs = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    st = {}
    st_ptrs = None
    if stages:
        st_ptrs = (ctypes.c_void_p * (3 * abi.NUM_STAGES))()
        for s in range(abi.NUM_STAGES):
            arr = np.zeros((3,) + stage_shape(frame, s), np.float32)
            st[abi.STAGE_NAMES[s]] = arr
            for c in range(3):
                st_ptrs[3 * s + c] = arr[c].ctypes.data
    rc = L.pikref_recon(ctypes.byref(fc), threads,
                        ctypes.cast(st_ptrs, ctypes.c_void_p) if st_ptrs else None,
                        ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikref_recon: %s" % L.pikref_last_error().decode())
    return out, st


def recon_passes(frame, extra, threads=0, kind="scalar"):
    """The reference reconstruction with AC passes after the frame's own
    (extra: list of (3, by, bx * 64) int16). Returns the linear (3, H, W)."""
    L = lib(kind)
    fc, keep = frame.to_c()
    extra = [np.ascontiguousarray(e, np.int16) for e in extra]
    ptrs = (ctypes.c_void_p * max(1, 3 * len(extra)))()
    for p, e in enumerate(extra):
        for c in range(3):
            ptrs[3 * p + c] = e[c].ctypes.data
    stride = extra[0].shape[2] if extra else 0
    out = np.zeros((3, frame.ysize, frame.xsize), np.float32)
    out_ptrs = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    rc = L.pikref_recon_passes(ctypes.byref(fc), threads, len(extra),
                               ctypes.cast(ptrs, ctypes.c_void_p), stride,
                               None, ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikref_recon_passes: %s"
                           % L.pikref_last_error().decode())
    return out


def recon_downsampled(frame, ds, threads=0, kind="scalar"):
    """The reference's downsampled decode (ds = 2, 4, 8): (3, ceil(H/ds),
    ceil(W/ds)) float32."""
    L = lib(kind)
    fc, keep = frame.to_c()
    oy, ox = -(-frame.ysize // ds), -(-frame.xsize // ds)
    out = np.zeros((3, oy, ox), np.float32)
    out_ptrs = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    rc = L.pikref_recon_downsampled(ctypes.byref(fc), threads, ds,
                                    ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikref_recon_downsampled: %s"
                           % L.pikref_last_error().decode())
    return out


def transform(raw, direction, block):
    """Reference TransformToPixels (direction 0) / TransformFromPixels (1)."""
    s = 4 if raw == abi.ACS_DCT32X32 else (2 if raw == abi.ACS_DCT16X16 else 1)
    inp = np.ascontiguousarray(block, np.float32).reshape(-1)
    assert inp.size == 64 * s * s
    out = np.zeros(64 * s * s, np.float32)
    lib().pikref_transform(raw, direction, inp.ctypes.data, out.ctypes.data)
    return out.reshape(8 * s, 8 * s)


def epf_weight(sigma, sad):
    return lib().pikref_epf_weight(sigma, sad)


def gaborish(strength, img):
    img = np.ascontiguousarray(img, np.float32)
    _, ys, xs = img.shape
    out = np.zeros_like(img)
    ip = (ctypes.c_void_p * 3)(*[img[c].ctypes.data for c in range(3)])
    op = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    rc = lib().pikref_gaborish(strength, ip, xs, ys, op)
    assert rc == 0
    return out


def opsin_to_linear(img):
    img = np.array(img, np.float32, copy=True)
    _, ys, xs = img.shape
    p = (ctypes.c_void_p * 3)(*[img[c].ctypes.data for c in range(3)])
    lib().pikref_opsin_to_linear(p, xs, ys)
    return img


def quantize(frame, opsin, kind="scalar"):
    """The reference encoder's quantized coefficients of the 8-padded XYB
    image `opsin` (3, by * 8, bx * 8) under the frame's side information
    (InitializeFrameEncCache + ComputeCoefficients per group, as
    RoundtripImage runs them). Returns (ac (3, by, bx * 64) int16,
    dc (3, by, bx) int16)."""
    L = lib(kind)
    fc, keep = frame.to_c()
    bx, by = frame.xsize_blocks, frame.ysize_blocks
    img = np.ascontiguousarray(opsin, np.float32)
    assert img.shape == (3, by * 8, bx * 8), img.shape
    ac = np.zeros((3, by, bx * 64), np.int16)
    dc = np.zeros((3, by, bx), np.int16)
    ip = (ctypes.c_void_p * 3)(*[img[c].ctypes.data for c in range(3)])
    ap = (ctypes.c_void_p * 3)(*[ac[c].ctypes.data for c in range(3)])
    dp = (ctypes.c_void_p * 3)(*[dc[c].ctypes.data for c in range(3)])
    rc = L.pikref_quantize(ctypes.byref(fc), ctypes.cast(ip, ctypes.c_void_p),
                           ctypes.cast(ap, ctypes.c_void_p),
                           ctypes.cast(dp, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikref_quantize: %s" % L.pikref_last_error().decode())
    return ac, dc


def _planes(img):
    img = np.ascontiguousarray(img, np.float32)
    return img, (ctypes.c_void_p * 3)(*[img[c].ctypes.data for c in range(3)])


def butteraugli(opsin, linear, distmap=False, kind="scalar"):
    """The reference Butteraugli distance of `linear` (3, H, W) from the
    image whose unpadded XYB planes are `opsin` (3, H, W), as the encoder's
    comparator computes it. Returns the distance, or (distance, map) with
    distmap=True."""
    L = lib(kind)
    L.pikref_butteraugli.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_size_t, ctypes.c_size_t,
                                     ctypes.c_void_p, ctypes.c_void_p]
    o, op = _planes(opsin)
    li, lp = _planes(linear)
    assert o.shape == li.shape, (o.shape, li.shape)
    _, ys, xs = o.shape
    dm = np.zeros((ys, xs), np.float32) if distmap else None
    d = ctypes.c_double()
    rc = L.pikref_butteraugli(ctypes.cast(op, ctypes.c_void_p),
                              ctypes.cast(lp, ctypes.c_void_p), xs, ys,
                              dm.ctypes.data if distmap else None,
                              ctypes.byref(d))
    if rc != 0:
        raise RuntimeError("pikref_butteraugli: %s"
                           % L.pikref_last_error().decode())
    return (d.value, dm) if distmap else d.value


def initial_quant_field(opsin, target, kind="scalar"):
    """The reference's InitialQuantField of the unpadded XYB image (3, H, W)
    at distance `target`: (ceil(H/8), ceil(W/8)) float32."""
    L = lib(kind)
    L.pikref_initial_quant_field.argtypes = [
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_float,
        ctypes.c_void_p]
    o, op = _planes(opsin)
    _, ys, xs = o.shape
    qf = np.zeros((div_ceil(ys, 8), div_ceil(xs, 8)), np.float32)
    rc = L.pikref_initial_quant_field(ctypes.cast(op, ctypes.c_void_p), xs,
                                      ys, target, qf.ctypes.data)
    if rc != 0:
        raise RuntimeError("pikref_initial_quant_field: %s"
                           % L.pikref_last_error().decode())
    return qf


def initial_quant_dc(target, kind="scalar"):
    L = lib(kind)
    L.pikref_initial_quant_dc.argtypes = [ctypes.c_float]
    L.pikref_initial_quant_dc.restype = ctypes.c_float
    return float(L.pikref_initial_quant_dc(target))


def color_correlation(opsin, kind="scalar"):
    """The reference's FindBestColorCorrelationMap of the padded XYB image
    (3, H8, W8): (ytox (ty, tx) int32, ytob, ytox_dc, ytob_dc)."""
    L = lib(kind)
    L.pikref_color_correlation.argtypes = [
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p]
    o, op = _planes(opsin)
    _, ys, xs = o.shape
    tx, ty = div_ceil(xs, 64), div_ceil(ys, 64)
    ytox = np.zeros((ty, tx), np.int32)
    ytob = np.zeros((ty, tx), np.int32)
    dc = np.zeros(2, np.int32)
    rc = L.pikref_color_correlation(ctypes.cast(op, ctypes.c_void_p), xs, ys,
                                    ytox.ctypes.data, ytob.ctypes.data,
                                    dc.ctypes.data)
    if rc != 0:
        raise RuntimeError("pikref_color_correlation: %s"
                           % L.pikref_last_error().decode())
    return ytox, ytob, int(dc[0]), int(dc[1])


def encoded_bytes(frame, kind="scalar"):
    """(AC bytes, DC + control-field bytes) of the frame's quantized data
    with the reference's entropy coder."""
    L = lib(kind)
    L.pikref_encoded_bytes.argtypes = [ctypes.POINTER(abi.FrameC),
                                       ctypes.c_void_p, ctypes.c_void_p]
    fc, keep = frame.to_c()
    a, d = ctypes.c_size_t(), ctypes.c_size_t()
    rc = L.pikref_encoded_bytes(ctypes.byref(fc), ctypes.byref(a),
                                ctypes.byref(d))
    if rc != 0:
        raise RuntimeError("pikref_encoded_bytes: %s"
                           % L.pikref_last_error().decode())
    return a.value, d.value


def find_best_quantization(frame, opsin, opsin_orig, init, target,
                           rounds=12, threads=1, kind="scalar"):
    """The reference encoder's Butteraugli-driven quant-field search over
    the frame's side information (restated over the reference's own
    functions). opsin: padded XYB (3, by*8, bx*8) the encoder quantizes;
    opsin_orig: unpadded XYB (3, H, W); init: initial float field (by, bx).
    Returns dict(quant_field (by, bx) int32, global_scale, quant_dc,
    distance, seconds)."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_find_best_quantization.argtypes = [
        ctypes.POINTER(abi.FrameC), vp, vp, vp, ctypes.c_float, ctypes.c_int,
        ctypes.c_int, vp, vp, vp, vp]
    fc, keep = frame.to_c()
    o, op = _planes(opsin)
    oo, oop = _planes(opsin_orig)
    init = np.ascontiguousarray(init, np.float32)
    by, bx = frame.ysize_blocks, frame.xsize_blocks
    assert init.shape == (by, bx)
    qf = np.zeros((by, bx), np.int32)
    sc = np.zeros(2, np.int32)
    d, sec = ctypes.c_double(), ctypes.c_double()
    rc = L.pikref_find_best_quantization(
        ctypes.byref(fc), ctypes.cast(op, vp), ctypes.cast(oop, vp),
        init.ctypes.data, target, rounds, threads, qf.ctypes.data,
        sc.ctypes.data, ctypes.byref(d), ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError("pikref_find_best_quantization: %s"
                           % L.pikref_last_error().decode())
    return dict(quant_field=qf, global_scale=int(sc[0]), quant_dc=int(sc[1]),
                distance=d.value, seconds=sec.value)


def encode_ac_groups(frame, noise=None, fast_mode=False, kind="scalar"):
    """The frame's AC as the reference encoder writes it, one bitstream per
    512x512 group (EncodeToBitstream). Returns a list of bytes objects."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_encode_ac_groups.argtypes = [
        ctypes.POINTER(abi.FrameC), vp, ctypes.c_int, vp, ctypes.c_size_t,
        vp, vp]
    fc, keep = frame.to_c()
    ng = (-(-frame.xsize // 512)) * (-(-frame.ysize // 512))
    sizes = np.zeros(ng, np.uintp)
    total = ctypes.c_size_t()
    nz = None if noise is None else np.asarray(noise, np.float32)
    nzp = None if nz is None else nz.ctypes.data
    rc = L.pikref_encode_ac_groups(ctypes.byref(fc), nzp, int(fast_mode),
                                   None, 0, sizes.ctypes.data,
                                   ctypes.byref(total))
    if rc != 0:
        raise RuntimeError("pikref_encode_ac_groups: %s"
                           % L.pikref_last_error().decode())
    buf = np.zeros(max(1, total.value), np.uint8)
    rc = L.pikref_encode_ac_groups(ctypes.byref(fc), nzp, int(fast_mode),
                                   buf.ctypes.data, buf.size,
                                   sizes.ctypes.data, ctypes.byref(total))
    if rc != 0:
        raise RuntimeError("pikref_encode_ac_groups: %s"
                           % L.pikref_last_error().decode())
    out, pos = [], 0
    for s in sizes:
        out.append(buf[pos:pos + int(s)].tobytes())
        pos += int(s)
    return out


def decode_ac_groups(frame, groups, check_dequant=True, kind="scalar",
                     out=None):
    """Decodes group bitstreams with the reference decoder's functions (and,
    with check_dequant, checks DecodeFromBitstream<true> agrees). Returns
    (ac int16 (3, by, bx*64), noise (groups, 3), consumed bytes per group)."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_decode_ac_groups.argtypes = [
        ctypes.POINTER(abi.FrameC), vp, vp, vp, ctypes.c_size_t, vp, vp,
        ctypes.c_int]
    fc, keep = frame.to_c()
    by, bx = frame.ysize_blocks, frame.xsize_blocks
    data = np.frombuffer(b"".join(groups) + b"\0", np.uint8)
    sizes = np.array([len(g) for g in groups], np.uintp)
    ac = np.zeros((3, by, bx * 64), np.int16) if out is None else out
    assert ac.shape == (3, by, bx * 64) and ac.dtype == np.int16
    ptrs = (vp * 3)(*[ac[c].ctypes.data for c in range(3)])
    noise = np.zeros((len(groups), 3), np.float32)
    used = np.zeros(len(groups), np.uintp)
    rc = L.pikref_decode_ac_groups(ctypes.byref(fc), data.ctypes.data,
                                   sizes.ctypes.data, ctypes.cast(ptrs, vp),
                                   bx * 64, noise.ctypes.data,
                                   used.ctypes.data, int(check_dequant))
    if rc != 0:
        raise RuntimeError("pikref_decode_ac_groups: %s"
                           % L.pikref_last_error().decode())
    return ac, noise, used.astype(np.int64)


def encode_dc_groups(frame, kind="scalar"):
    """The frame's DC groups section (size table + groups) as the reference
    encoder writes it: quantized DC, AC strategy, quant field, sigma LUT
    ids. Returns bytes."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_encode_dc_groups.argtypes = [ctypes.POINTER(abi.FrameC), vp,
                                          ctypes.c_size_t, vp]
    fc, keep = frame.to_c()
    total = ctypes.c_size_t()
    if L.pikref_encode_dc_groups(ctypes.byref(fc), None, 0,
                                 ctypes.byref(total)) != 0:
        raise RuntimeError("pikref_encode_dc_groups: %s"
                           % L.pikref_last_error().decode())
    buf = np.zeros(max(1, total.value), np.uint8)
    if L.pikref_encode_dc_groups(ctypes.byref(fc), buf.ctypes.data, buf.size,
                                 ctypes.byref(total)) != 0:
        raise RuntimeError("pikref_encode_dc_groups: %s"
                           % L.pikref_last_error().decode())
    return buf[:total.value].tobytes()


def decode_dc_groups(xsize, ysize, data, kind="scalar"):
    """The reference decoder's reading of a DC groups section: dict(dc
    (3, by, bx) int16, ac_strategy (by, bx) uint8, quant_field int32,
    sigma_lut_ids uint8, consumed)."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_decode_dc_groups.argtypes = [
        ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_size_t, vp, vp, vp, vp,
        vp]
    by, bx = -(-ysize // 8), -(-xsize // 8)
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    dc = np.zeros((3, by, bx), np.int16)
    ptrs = (vp * 3)(*[dc[c].ctypes.data for c in range(3)])
    acs = np.zeros((by, bx), np.uint8)
    qf = np.zeros((by, bx), np.int32)
    lut = np.zeros((by, bx), np.uint8)
    used = ctypes.c_size_t()
    rc = L.pikref_decode_dc_groups(xsize, ysize, buf.ctypes.data, len(data),
                                   ctypes.cast(ptrs, vp), acs.ctypes.data,
                                   qf.ctypes.data, lut.ctypes.data,
                                   ctypes.byref(used))
    if rc != 0:
        raise RuntimeError("pikref_decode_dc_groups: %s"
                           % L.pikref_last_error().decode())
    return dict(dc=dc, ac_strategy=acs, quant_field=qf, sigma_lut_ids=lut,
                consumed=used.value)


# Bits of encode_frame's `variant` (pikref_encode_frame).
VAR_FRAME_INFO = 1
VAR_GROUP_HEADERS = 2
VAR_STATIC_DOTS = 4
VAR_FAST_AC = 8
VAR_DOWNSAMPLING = 16
VAR_COPY_TABLE = 32
VAR_EXTENSIONS = 64


def encode_frame(frame, extra_ac=(), noise=None, variant=0, kind="scalar"):
    """One whole frame (FrameHeader to the last pass's groups) written with
    the reference's encoders: pass 0 carries frame.ac, pass i > 0
    extra_ac[i - 1] ((3, by, bx*64) int16). noise: (passes, groups, 3) or
    None. Returns bytes."""
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_encode_frame.argtypes = [ctypes.POINTER(abi.FrameC), ctypes.c_int,
                                      vp, vp, ctypes.c_int, vp,
                                      ctypes.c_size_t, vp]
    fc, keep = frame.to_c()
    extra = [np.ascontiguousarray(a, np.int16) for a in extra_ac]
    for a in extra:
        assert a.shape == frame.ac.shape and fc.ac_stride == a.shape[2]
    ptrs = (vp * max(1, 3 * len(extra)))(
        *[a[c].ctypes.data for a in extra for c in range(3)])
    nz = None if noise is None else np.ascontiguousarray(noise, np.float32)
    npp = None if nz is None else nz.ctypes.data
    total = ctypes.c_size_t()
    args = (ctypes.byref(fc), 1 + len(extra), ctypes.cast(ptrs, vp), npp,
            int(variant))
    if L.pikref_encode_frame(*args, None, 0, ctypes.byref(total)) != 0:
        raise RuntimeError("pikref_encode_frame: %s"
                           % L.pikref_last_error().decode())
    buf = np.zeros(max(1, total.value), np.uint8)
    if L.pikref_encode_frame(*args, buf.ctypes.data, buf.size,
                             ctypes.byref(total)) != 0:
        raise RuntimeError("pikref_encode_frame: %s"
                           % L.pikref_last_error().decode())
    return buf[:total.value].tobytes()


FRAME_HEADER_FIELDS = ("flags", "gaborish", "predict_lf", "predict_hf",
                       "adaptive_reconstruction", "epf_enable_adaptive",
                       "epf_sigma", "epf_use_sharpened", "num_passes",
                       "global_scale", "quant_dc", "ytox_dc", "ytob_dc",
                       "num_table_sets", "num_dots", "frame_size")


def decode_frame(data, xsize, ysize, max_passes=16, max_dots=4096,
                 kind="scalar", reuse=None):
    """The reference decoders' reading of one frame: dict with the header
    fields (FRAME_HEADER_FIELDS), dc, ac_strategy, quant_field,
    sigma_lut_ids, ytox_map, ytob_map, dequant_cf, dequant_map, dots
    (structured, frame.DOT_DTYPE), passes (list of (3, by, bx*64) int16),
    noise (passes, groups, 3) and consumed. Raises RuntimeError where the
    reference fails. `reuse` (a dict, may be empty) keeps the output
    buffers between calls of the same sizes; the result then views them."""
    from pik_amd.frame import DOT_DTYPE
    L = lib(kind)
    vp = ctypes.c_void_p
    L.pikref_decode_frame.argtypes = [
        vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, vp,
        vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_size_t, vp, vp]
    by, bx = -(-ysize // 8), -(-xsize // 8)
    ty, tx = -(-ysize // 64), -(-xsize // 64)
    ng = (-(-xsize // 512)) * (-(-ysize // 512))
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8)
    key = (xsize, ysize, max_passes, max_dots)
    if reuse is None or reuse.get("key") != key:
        b = dict(key=key, hdr=np.zeros(16, np.int32),
                 dc=np.zeros((3, by, bx), np.int16),
                 acs=np.zeros((by, bx), np.uint8),
                 qf=np.zeros((by, bx), np.int32),
                 lut=np.zeros((by, bx), np.uint8),
                 ytox=np.zeros((ty, tx), np.int32),
                 ytob=np.zeros((ty, tx), np.int32),
                 dcf=np.zeros((ty, tx), np.uint8),
                 dqmap=np.zeros((16, 256), np.uint8),
                 dots=np.zeros(max_dots, DOT_DTYPE),
                 ac=np.zeros((max_passes, 3, by, bx * 64), np.int16),
                 noise=np.zeros((max_passes, ng, 3), np.float32))
        if reuse is not None:
            reuse.clear()
            reuse.update(b)
    else:
        b = reuse
    hdr, dc, acs, qf, lut = b["hdr"], b["dc"], b["acs"], b["qf"], b["lut"]
    ytox, ytob, dcf, dqmap = b["ytox"], b["ytob"], b["dcf"], b["dqmap"]
    dots, ac, noise = b["dots"], b["ac"], b["noise"]
    dptr = (vp * 3)(*[dc[c].ctypes.data for c in range(3)])
    used = ctypes.c_size_t()
    rc = L.pikref_decode_frame(
        buf.ctypes.data, len(data), xsize, ysize, hdr.ctypes.data,
        ctypes.cast(dptr, vp), acs.ctypes.data, qf.ctypes.data,
        lut.ctypes.data, ytox.ctypes.data, ytob.ctypes.data, dcf.ctypes.data,
        dqmap.ctypes.data, dots.ctypes.data, max_dots, ac.ctypes.data,
        max_passes, noise.ctypes.data, ctypes.byref(used))
    if rc != 0:
        raise RuntimeError("pikref_decode_frame: %s"
                           % L.pikref_last_error().decode())
    out = dict(zip(FRAME_HEADER_FIELDS, (int(v) for v in hdr)))
    n = out["num_passes"]
    if out["num_dots"] > max_dots:
        raise RuntimeError("more dots than max_dots")
    out.update(dc=dc, ac_strategy=acs, quant_field=qf, sigma_lut_ids=lut,
               ytox_map=ytox, ytob_map=ytob, dequant_cf=dcf,
               dequant_map=dqmap, dots=dots[:out["num_dots"]].copy(),
               passes=[ac[i] for i in range(n)], noise=noise[:n],
               consumed=int(used.value))
    return out
