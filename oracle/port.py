"""ctypes binding of the oracle's C restatement (pik_oracle.c, built by
`make -C oracle port` into oracle/_build/libpikoracle.so). Test
infrastructure: a checker, or the CPU baseline timed beside the GPU, never
part of the product."""
import ctypes
import os

import numpy as np

from pik_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libpikoracle.so")

_lib = None


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle port not built: %s" % LIB)
        L = ctypes.CDLL(LIB)
        L.pikoracle_recon.argtypes = [ctypes.POINTER(abi.FrameC),
                                      ctypes.c_void_p, ctypes.c_void_p]
        L.pikoracle_recon_passes.argtypes = [
            ctypes.POINTER(abi.FrameC), ctypes.c_int, ctypes.c_void_p,
            ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.pikoracle_recon_downsampled.argtypes = [
            ctypes.POINTER(abi.FrameC), ctypes.c_int, ctypes.c_void_p]
        L.pikoracle_last_error.restype = ctypes.c_char_p
        L.pikoracle_quantize.argtypes = [ctypes.POINTER(abi.FrameC),
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
        L.pikoracle_block_transform.argtypes = [ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p,
                                                ctypes.c_void_p]
        _lib = L
    return _lib


def stage_shape(frame, stage):
    bx, by = frame.xsize_blocks, frame.ysize_blocks
    if stage in (abi.STAGE_DC_DEQUANT, abi.STAGE_DC):
        return (by, bx)
    if stage == abi.STAGE_COEFFS:
        return (by, bx * 64)
    if stage == abi.STAGE_LINEAR:
        return (frame.ysize, frame.xsize)
    return (by * 8, bx * 8)


def recon(frame, stages=True):
    """Returns (linear (3, H, W) float32, dict stage name -> array)."""
    L = lib()
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
    rc = L.pikoracle_recon(ctypes.byref(fc),
                           ctypes.cast(st_ptrs, ctypes.c_void_p)
                           if st_ptrs else None,
                           ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikoracle_recon: %s"
                           % L.pikoracle_last_error().decode())
    return out, st


def recon_passes(frame, extra):
    """Reconstruction with AC passes after the frame's own: extra is a list
    of (3, by, bx * 64) int16 arrays. Returns the linear (3, H, W)."""
    L = lib()
    fc, keep = frame.to_c()
    extra = [np.ascontiguousarray(e, np.int16) for e in extra]
    ptrs = (ctypes.c_void_p * max(1, 3 * len(extra)))()
    for p, e in enumerate(extra):
        for c in range(3):
            ptrs[3 * p + c] = e[c].ctypes.data
    stride = extra[0].shape[2] if extra else 0
    out = np.zeros((3, frame.ysize, frame.xsize), np.float32)
    out_ptrs = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    rc = L.pikoracle_recon_passes(ctypes.byref(fc), len(extra),
                                  ctypes.cast(ptrs, ctypes.c_void_p), stride,
                                  None, ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikoracle_recon_passes: %s"
                           % L.pikoracle_last_error().decode())
    return out


def recon_downsampled(frame, ds):
    """Downsampled decode (ds = 2, 4, 8): (3, ceil(H/ds), ceil(W/ds))
    float32."""
    L = lib()
    fc, keep = frame.to_c()
    oy, ox = -(-frame.ysize // ds), -(-frame.xsize // ds)
    out = np.zeros((3, oy, ox), np.float32)
    out_ptrs = (ctypes.c_void_p * 3)(*[out[c].ctypes.data for c in range(3)])
    rc = L.pikoracle_recon_downsampled(ctypes.byref(fc), ds,
                                       ctypes.cast(out_ptrs, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikoracle_recon_downsampled: %s"
                           % L.pikoracle_last_error().decode())
    return out


def transform(raw, direction, block):
    """TransformToPixels (direction 0) / TransformFromPixels (1)."""
    n = block.shape[0]
    inp = np.ascontiguousarray(block, np.float32).reshape(-1)
    out = np.zeros(n * n, np.float32)
    rc = lib().pikoracle_block_transform(raw, direction, inp.ctypes.data,
                                         out.ctypes.data)
    assert rc == 0
    return out.reshape(n, n)


def quantize(frame, opsin):
    """The C restatement of the encoder's quantization (frames without
    predictions): (ac (3, by, bx * 64) int16, dc (3, by, bx) int16) of the
    8-padded XYB image `opsin` (3, by * 8, bx * 8)."""
    L = lib()
    fc, keep = frame.to_c()
    bx, by = frame.xsize_blocks, frame.ysize_blocks
    img = np.ascontiguousarray(opsin, np.float32)
    assert img.shape == (3, by * 8, bx * 8), img.shape
    ac = np.zeros((3, by, bx * 64), np.int16)
    dc = np.zeros((3, by, bx), np.int16)
    ip = (ctypes.c_void_p * 3)(*[img[c].ctypes.data for c in range(3)])
    ap = (ctypes.c_void_p * 3)(*[ac[c].ctypes.data for c in range(3)])
    dp = (ctypes.c_void_p * 3)(*[dc[c].ctypes.data for c in range(3)])
    rc = L.pikoracle_quantize(ctypes.byref(fc), ctypes.cast(ip, ctypes.c_void_p),
                              ctypes.cast(ap, ctypes.c_void_p),
                              ctypes.cast(dp, ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError("pikoracle_quantize: %s"
                           % L.pikoracle_last_error().decode())
    return ac, dc
