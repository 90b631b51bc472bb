"""Python host wrapper over the C ABI of libpikgpu.so (include/pikgpu.h).

Mirrors the reference's decode-side entry points for this path: one
`Reconstructor` plays the part of the FrameDecCache/GroupDecCache owner, and
`reconstruct(frame)` is the drop-in for the sequence DequantImageAC ->
ReconOpsinImage -> FinalizeFrameDecoding (pik/compressed_image.h:83-110).
Errors raise PikGpuError with the library's message (the reference returns a
false Status).
"""
import ctypes

import numpy as np

from . import abi


class PikGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pikgpu error %d: %s" % (code, msg))
        self.code = code


def validate(frame):
    """Host-only frame check (pikgpu_validate_frame). Raises PikGpuError
    with the library's reason when the frame is not decodable."""
    lib = abi.load()
    fc, keep = frame.to_c()
    msg = ctypes.create_string_buffer(256)
    rc = lib.pikgpu_validate_frame(ctypes.byref(fc), msg, len(msg))
    if rc != abi.OK:
        raise PikGpuError(rc, msg.value.decode())


class Reconstructor:
    def __init__(self, device=0):
        self._lib = abi.load()
        cfg = abi.ConfigC(abi.ABI_VERSION, 0)
        self._ctx = self._lib.pikgpu_create(device, ctypes.byref(cfg))
        if not self._ctx:
            raise PikGpuError(abi.EHIP, "pikgpu_create failed on device %d"
                              % device)
        self._frame = None
        self._keep = None

    def close(self):
        if self._ctx:
            self._lib.pikgpu_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != abi.OK:
            msg = self._lib.pikgpu_last_error(self._ctx)
            raise PikGpuError(rc, msg.decode() if msg else "")

    @property
    def stream(self):
        return self._lib.pikgpu_stream(self._ctx)

    def upload(self, frame):
        fc, keep = frame.to_c()
        self._check(self._lib.pikgpu_frame_upload(self._ctx, ctypes.byref(fc)))
        self._frame = frame
        self._keep = (fc, keep)

    def run(self, keep_stages=False):
        self._check(self._lib.pikgpu_run(self._ctx, int(bool(keep_stages))))

    def sync(self):
        self._check(self._lib.pikgpu_sync(self._ctx))

    def download(self):
        f = self._frame
        out = np.empty((3, f.ysize, f.xsize), np.float32)
        planes = (abi.c_float_p * 3)(
            *[out[c].ctypes.data_as(abi.c_float_p) for c in range(3)])
        self._check(self._lib.pikgpu_download(self._ctx, planes,
                                              f.xsize * 4))
        return out

    def set_output(self, outputs):
        """abi.OUTPUT_LINEAR | abi.OUTPUT_SRGB8 for the following runs."""
        self._check(self._lib.pikgpu_set_output(self._ctx, int(outputs)))

    def download_srgb8(self):
        """(ysize, xsize, 3) uint8 sRGB of the last run."""
        f = self._frame
        out = np.empty((f.ysize, f.xsize, 3), np.uint8)
        self._check(self._lib.pikgpu_download_srgb8(
            self._ctx, out.ctypes.data, f.xsize * 3))
        return out

    def reconstruct(self, frame):
        """Upload + run + download; returns (3, ysize, xsize) float32."""
        self.upload(frame)
        self.run(False)
        return self.download()

    def stage(self, stage):
        xs, ys = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.pikgpu_stage_size(self._ctx, stage,
                                                ctypes.byref(xs),
                                                ctypes.byref(ys)))
        out = np.empty((3, ys.value, xs.value), np.float32)
        planes = (abi.c_float_p * 3)(
            *[out[c].ctypes.data_as(abi.c_float_p) for c in range(3)])
        self._check(self._lib.pikgpu_stage_dump(self._ctx, stage, planes,
                                                xs.value * 4))
        return out

    def timings(self):
        names = (ctypes.c_char_p * 16)()
        ms = (ctypes.c_float * 16)()
        n = self._lib.pikgpu_last_timings(self._ctx, names, ms, 16)
        return {names[i].decode(): float(ms[i]) for i in range(n)}


class BatchReconstructor:
    """Many frames of uneven sizes reconstructed by one graph launch
    (pikgpu_batch_*): upload once, run, download frame by frame."""

    def __init__(self, device=0):
        self._lib = abi.load()
        cfg = abi.ConfigC(abi.ABI_VERSION, 0)
        self._b = self._lib.pikgpu_batch_create(device, ctypes.byref(cfg))
        if not self._b:
            raise PikGpuError(abi.EHIP, "pikgpu_batch_create failed on "
                              "device %d" % device)
        self._frames = []
        self._keep = None

    def close(self):
        if self._b:
            self._lib.pikgpu_batch_destroy(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != abi.OK:
            msg = self._lib.pikgpu_batch_last_error(self._b)
            raise PikGpuError(rc, msg.decode() if msg else "")

    def upload(self, frames):
        frames = list(frames)
        arr = (abi.FrameC * len(frames))()
        keep = []
        for i, fr in enumerate(frames):
            fc, k = fr.to_c()
            arr[i] = fc
            keep.append(k)
        self._check(self._lib.pikgpu_batch_upload(
            self._b, ctypes.cast(arr, ctypes.c_void_p), len(frames)))
        self._frames = frames
        self._keep = (arr, keep)

    def run(self):
        self._check(self._lib.pikgpu_batch_run(self._b))

    def sync(self):
        self._check(self._lib.pikgpu_batch_sync(self._b))

    @property
    def last_ms(self):
        return float(self._lib.pikgpu_batch_last_ms(self._b))

    def download(self, i):
        f = self._frames[i]
        out = np.empty((3, f.ysize, f.xsize), np.float32)
        planes = (abi.c_float_p * 3)(
            *[out[c].ctypes.data_as(abi.c_float_p) for c in range(3)])
        self._check(self._lib.pikgpu_batch_download(self._b, i, planes,
                                                    f.xsize * 4))
        return out

    def reconstruct(self, frames):
        self.upload(frames)
        self.run()
        return [self.download(i) for i in range(len(self._frames))]
