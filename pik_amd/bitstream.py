"""Host entropy decoding of AC group bitstreams (pikgpu_decode_ac_pass).

The reference decodes each 512x512 group's pass data in
DecodeFromBitstream (pik/compressed_image.cc:719-813) and dequantizes per
tile as it goes; here the int16 coefficients of every group land in one
frame-wide array (the layout of Frame.ac) that the GPU reconstruction then
takes whole. The decoding runs in libpikgpu.so on host threads.
"""
import ctypes

import numpy as np

from . import abi
from .runtime import PikGpuError


def num_groups(xsize, ysize):
    return (-(-xsize // 512)) * (-(-ysize // 512))


def decode_ac_pass(xsize, ysize, ac_strategy, groups, threads=0, out=None):
    """Decodes one pass: `groups` holds each group's bytes (row-major over
    the 512x512 groups), `ac_strategy` the (by, bx) uint8 strategy bytes.
    Returns dict(ac=(3, by, bx*64) int16, noise=(groups, 3) float32,
    consumed=(groups,) bytes read). Raises PikGpuError (code abi.EDATA for
    an invalid bitstream)."""
    lib = abi.load()
    by, bx = -(-ysize // 8), -(-xsize // 8)
    acs = np.ascontiguousarray(ac_strategy, np.uint8)
    if acs.shape != (by, bx):
        raise ValueError("ac_strategy must be (%d, %d)" % (by, bx))
    ng = num_groups(xsize, ysize)
    if len(groups) != ng:
        raise ValueError("%d groups given, the frame has %d"
                         % (len(groups), ng))
    if out is None:
        out = np.empty((3, by, bx * 64), np.int16)
    if out.shape != (3, by, bx * 64) or out.dtype != np.int16 or \
            not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous int16 (3, %d, %d)"
                         % (by, bx * 64))
    bufs = [np.frombuffer(g, np.uint8) if len(g) else np.zeros(1, np.uint8)
            for g in groups]
    ptrs = np.array([b.ctypes.data for b in bufs], np.uintp)
    sizes = np.array([len(g) for g in groups], np.uintp)
    noise = np.zeros((ng, 3), np.float32)
    used = np.zeros(ng, np.uintp)
    p = abi.AcPassC()
    p.xsize, p.ysize = xsize, ysize
    p.ac_strategy = acs.ctypes.data
    p.ac_strategy_stride = bx
    p.num_groups = ng
    p.group_data = ptrs.ctypes.data
    p.group_size = sizes.ctypes.data
    for c in range(3):
        p.ac[c] = out[c].ctypes.data
    p.ac_stride = bx * 64
    p.noise = noise.ctypes.data
    p.group_consumed = used.ctypes.data
    msg = ctypes.create_string_buffer(256)
    rc = lib.pikgpu_decode_ac_pass(ctypes.byref(p), int(threads), msg,
                                   len(msg))
    if rc != abi.OK:
        raise PikGpuError(rc, msg.value.decode())
    return dict(ac=out, noise=noise, consumed=used.astype(np.int64))


def decode_dc_groups(xsize, ysize, data, threads=0):
    """Decodes a frame's DC groups section (size table + 256x256-block
    groups): dict(dc=(3, by, bx) int16 quantized DC, ac_strategy=(by, bx)
    uint8, quant_field=(by, bx) int32, sigma_lut_ids=(by, bx) uint8,
    consumed=bytes of the section). Raises PikGpuError (abi.EDATA for an
    invalid section)."""
    lib = abi.load()
    by, bx = -(-ysize // 8), -(-xsize // 8)
    raw = np.frombuffer(bytes(data), np.uint8) if len(data) else \
        np.zeros(1, np.uint8)
    dc = np.zeros((3, by, bx), np.int16)
    acs = np.zeros((by, bx), np.uint8)
    qf = np.zeros((by, bx), np.int32)
    lut = np.zeros((by, bx), np.uint8)
    s = abi.DcSectionC()
    s.xsize, s.ysize = xsize, ysize
    s.data = raw.ctypes.data
    s.size = len(data)
    for c in range(3):
        s.dc[c] = dc[c].ctypes.data
    s.dc_stride = bx
    s.ac_strategy = acs.ctypes.data
    s.ac_strategy_stride = bx
    s.quant_field = qf.ctypes.data
    s.quant_field_stride = bx
    s.sigma_lut_ids = lut.ctypes.data
    s.sigma_lut_stride = bx
    msg = ctypes.create_string_buffer(256)
    rc = lib.pikgpu_decode_dc_groups(ctypes.byref(s), int(threads), msg,
                                     len(msg))
    if rc != abi.OK:
        raise PikGpuError(rc, msg.value.decode())
    return dict(dc=dc, ac_strategy=acs, quant_field=qf, sigma_lut_ids=lut,
                consumed=int(s.consumed))


def parse_frame(data, xsize, ysize, threads=0):
    """Parses one whole frame (FrameHeader to the end of its last pass;
    `data` may run on past it) of an xsize x ysize image. Returns
    (frame, info): frame a Frame holding pass 0's AC, info a dict with
    passes (list of (3, by, bx*64) int16, pass 0 first), noise
    (passes, groups, 3), frame_size, epf_sigma, downsampling (list of
    (factor, passes)) and consumed (bytes of the frame). Raises PikGpuError
    (abi.EDATA for an invalid frame, abi.EUNSUPPORTED for one outside this
    decoder)."""
    from .frame import DOT_DTYPE, Frame
    lib = abi.load()
    raw = np.frombuffer(bytes(data), np.uint8) if len(data) else \
        np.zeros(1, np.uint8)
    handle = ctypes.c_void_p()
    msg = ctypes.create_string_buffer(256)
    rc = lib.pikgpu_parse_frame(raw.ctypes.data, len(data), xsize, ysize,
                                int(threads), ctypes.byref(handle), msg,
                                len(msg))
    if rc != abi.OK:
        raise PikGpuError(rc, msg.value.decode())
    try:
        fc = abi.FrameC()
        info = abi.ParsedInfoC()
        lib.pikgpu_parsed_get(handle, ctypes.byref(fc), ctypes.byref(info))
        by, bx = -(-ysize // 8), -(-xsize // 8)
        ty, tx = -(-ysize // 64), -(-xsize // 64)

        def arr(ptr, dtype, shape):
            n = int(np.prod(shape))
            if n == 0:
                return np.zeros(shape, dtype)
            ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)),
                                      (n,))
            return a.reshape(shape).copy()

        passes = []
        for i in range(info.num_passes):
            planes = (ctypes.c_void_p * 3)()
            stride = ctypes.c_size_t()
            lib.pikgpu_parsed_pass_ac(handle, i, planes, ctypes.byref(stride))
            passes.append(np.stack([arr(planes[c], np.int16, (by, bx * 64))
                                    for c in range(3)]))
        dots = np.zeros(fc.num_dots, DOT_DTYPE)
        if fc.num_dots:
            ctypes.memmove(dots.ctypes.data, fc.dots,
                           fc.num_dots * DOT_DTYPE.itemsize)
        frame = Frame(
            xsize=xsize, ysize=ysize, ac=passes[0],
            dc=np.stack([arr(fc.dc[c], np.int16, (by, bx)) for c in range(3)]),
            quant_field=arr(fc.quant_field, np.int32, (by, bx)),
            ac_strategy=arr(fc.ac_strategy, np.uint8, (by, bx)),
            sigma_lut_ids=arr(fc.sigma_lut_ids, np.uint8, (by, bx)),
            dequant_cf=arr(fc.dequant_cf, np.uint8, (ty, tx)),
            ytox_map=arr(fc.ytox_map, np.int32, (ty, tx)),
            ytob_map=arr(fc.ytob_map, np.int32, (ty, tx)),
            dequant_map=arr(fc.dequant_map, np.uint8, (16, 256)),
            dequant_table=arr(fc.dequant_table, np.float32,
                              (fc.dequant_table_size,)),
            dequant_offsets=arr(fc.dequant_offsets, np.uint32,
                                (fc.num_dequant_offsets,)),
            dots=dots, global_scale=fc.global_scale, quant_dc=fc.quant_dc,
            ytox_dc=fc.ytox_dc, ytob_dc=fc.ytob_dc, gaborish=fc.gaborish,
            predict_lf=fc.predict_lf, predict_hf=fc.predict_hf,
            adaptive_reconstruction=fc.adaptive_reconstruction,
            epf_enable_adaptive=fc.epf_enable_adaptive,
            epf_use_sharpened=fc.epf_use_sharpened, grayscale=fc.grayscale,
            flags=fc.flags)
        ng = info.num_groups
        ds = arr(info.downsampling, np.uint32, (2 * info.num_downsampling,))
        out = dict(passes=passes,
                   noise=arr(info.noise, np.float32,
                             (info.num_passes, ng, 3)),
                   frame_size=int(info.frame_size),
                   epf_sigma=int(info.epf_sigma),
                   downsampling=[(int(ds[2 * i]), int(ds[2 * i + 1]))
                                 for i in range(info.num_downsampling)],
                   consumed=int(info.consumed))
        return frame, out
    finally:
        lib.pikgpu_parsed_free(handle)
