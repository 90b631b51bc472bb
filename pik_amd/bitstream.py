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
