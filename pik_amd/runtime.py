"""Python host wrapper over the C ABI of libpikgpu.so (include/pikgpu.h).

Mirrors the reference's decode-side entry points for this path: one
`Reconstructor` plays the part of the FrameDecCache/GroupDecCache owner, and
`reconstruct(frame)` is the drop-in for the sequence DequantImageAC ->
ReconOpsinImage -> FinalizeFrameDecoding (pik/compressed_image.h:83-110).
Errors raise PikGpuError with the library's message (the reference returns a
false Status).
"""
import ctypes
import types

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
        self._ds = 1

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

    def decode_frame(self, data, xsize, ysize, threads=0):
        """Parses one whole frame of an xsize x ysize image on the host and
        reconstructs it (pikgpu_decode_frame: parse, upload, every further
        pass, run); download() then returns the image. Returns the frame's
        length in bytes."""
        raw = np.frombuffer(bytes(data), np.uint8) if len(data) else \
            np.zeros(1, np.uint8)
        used = ctypes.c_size_t()
        self._frame = None
        self._check(self._lib.pikgpu_decode_frame(
            self._ctx, raw.ctypes.data, len(data), xsize, ysize,
            int(threads), ctypes.byref(used)))
        self._frame = types.SimpleNamespace(
            xsize=xsize, ysize=ysize, xsize_blocks=-(-xsize // 8),
            ysize_blocks=-(-ysize // 8))
        self._keep = None
        return int(used.value)

    def run(self, keep_stages=False):
        self._check(self._lib.pikgpu_run(self._ctx, int(bool(keep_stages))))

    def sync(self):
        self._check(self._lib.pikgpu_sync(self._ctx))

    def _out_size(self):
        f, ds = self._frame, self._ds
        return -(-f.xsize // ds), -(-f.ysize // ds)

    def download(self, out=None):
        """(3, ysize, xsize) float32 (ceil(size / factor) when
        downsampled). `out`, if given, receives it: a (3, ysize, xsize)
        float32 array whose rows may be padded (a view into wider rows, as
        the reference's Image3F holds them), written in place."""
        xs, ys = self._out_size()
        if out is None:
            out = np.empty((3, ys, xs), np.float32)
        if (out.dtype != np.float32 or out.shape != (3, ys, xs) or
                out.strides[2] != 4 or out.strides[1] < 4 * xs):
            raise ValueError("out must be (3, %d, %d) float32 with unit-"
                             "stride rows" % (ys, xs))
        planes = (abi.c_float_p * 3)(
            *[out[c].ctypes.data_as(abi.c_float_p) for c in range(3)])
        self._check(self._lib.pikgpu_download(self._ctx, planes,
                                              out.strides[1]))
        return out

    def add_pass(self, ac):
        """Adds an AC pass (3 planes of int16, the frame's AC layout) to the
        uploaded frame (pikgpu_frame_add_pass)."""
        f = self._frame
        if len(ac) != 3:
            raise ValueError("a pass has three planes")
        planes = [np.ascontiguousarray(a, np.int16) for a in ac]
        want = None if f is None else (f.ysize_blocks, f.xsize_blocks * 64)
        for a in planes:
            if want is not None and a.shape != want:
                raise ValueError("AC pass planes must be %s int16, got %s"
                                 % (want, a.shape))
        ptrs = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in planes])
        self._check(self._lib.pikgpu_frame_add_pass(
            self._ctx, ctypes.cast(ptrs, ctypes.c_void_p), planes[0].shape[1]))

    def quantize(self, opsin):
        """Encoder-side quantization of the 8-padded XYB image `opsin`
        (3, by * 8, bx * 8) under the uploaded frame's side information
        (pikgpu_quantize). The result replaces the frame's coefficients on
        the device, so run() then decodes the round trip. Returns
        (ac (3, by, bx * 64) int16, dc (3, by, bx) int16)."""
        f = self._frame
        bx, by = f.xsize_blocks, f.ysize_blocks
        img = np.ascontiguousarray(opsin, np.float32)
        if img.shape != (3, by * 8, bx * 8):
            raise ValueError("opsin must be (3, %d, %d)" % (by * 8, bx * 8))
        ac = np.zeros((3, by, bx * 64), np.int16)
        dc = np.zeros((3, by, bx), np.int16)
        vp = ctypes.c_void_p
        ip = (vp * 3)(*[img[c].ctypes.data for c in range(3)])
        ap = (vp * 3)(*[ac[c].ctypes.data for c in range(3)])
        dp = (vp * 3)(*[dc[c].ctypes.data for c in range(3)])
        self._check(self._lib.pikgpu_quantize(
            self._ctx, ctypes.cast(ip, vp), bx * 8, ctypes.cast(ap, vp),
            bx * 64, ctypes.cast(dp, vp), bx))
        return ac, dc

    def search_quant_field(self, opsin, target, steps=20, learning_rate=0.05,
                           beta1=0.9, beta2=0.999, eps=1e-8, lam=0.0,
                           qf_max=256, weight=None, trust=(0.0625, 16.0)):
        """Gradient quant-field search on the uploaded frame
        (pikgpu_search_quant_field). opsin: the padded XYB image the encoder
        quantizes (3, by * 8, bx * 8); target: the linear RGB original
        (3, ysize, xsize); weight: optional per-block weights (by, bx);
        trust: the field's bounds relative to its initial value.
        lam > 0 is the rate weight; lam <= 0 keeps the initial distortion
        level times -lam (0: times 1). The context then holds the frame
        quantized with the final field (run() decodes it). Returns a dict:
        quant_field (by, bx) int32, ac, dc (int16), loss and rate
        (steps + 1 float64 each), lam (the rate weight used)."""
        f = self._frame
        bx, by = f.xsize_blocks, f.ysize_blocks
        img = np.ascontiguousarray(opsin, np.float32)
        if img.shape != (3, by * 8, bx * 8):
            raise ValueError("opsin must be (3, %d, %d)" % (by * 8, bx * 8))
        tgt = np.ascontiguousarray(target, np.float32)
        if tgt.shape != (3, f.ysize, f.xsize):
            raise ValueError("target must be (3, %d, %d)" % (f.ysize, f.xsize))
        vp = ctypes.c_void_p
        prm = abi.SearchParamsC(int(steps), learning_rate, beta1, beta2, eps,
                                lam, int(qf_max), float(trust[0]),
                                float(trust[1]))
        io = abi.SearchIoC()
        io.opsin = (vp * 3)(*[img[c].ctypes.data for c in range(3)])
        io.opsin_stride = bx * 8
        io.target = (vp * 3)(*[tgt[c].ctypes.data for c in range(3)])
        io.target_stride = f.xsize
        if weight is not None:
            weight = np.ascontiguousarray(weight, np.float32)
            if weight.shape != (by, bx):
                raise ValueError("weight must be (%d, %d)" % (by, bx))
            io.weight = weight.ctypes.data
            io.weight_stride = bx
        out = {"quant_field": np.zeros((by, bx), np.int32),
               "ac": np.zeros((3, by, bx * 64), np.int16),
               "dc": np.zeros((3, by, bx), np.int16),
               "loss": np.zeros(steps + 1, np.float64),
               "rate": np.zeros(steps + 1, np.float64)}
        io.quant_field = out["quant_field"].ctypes.data
        io.quant_field_stride = bx
        io.ac = (vp * 3)(*[out["ac"][c].ctypes.data for c in range(3)])
        io.ac_stride = bx * 64
        io.dc = (vp * 3)(*[out["dc"][c].ctypes.data for c in range(3)])
        io.dc_stride = bx
        io.loss = out["loss"].ctypes.data
        io.rate = out["rate"].ctypes.data
        lam_used = ctypes.c_float()
        io.lambda_used = ctypes.addressof(lam_used)
        self._check(self._lib.pikgpu_search_quant_field(
            self._ctx, ctypes.byref(prm), ctypes.byref(io)))
        out["lam"] = lam_used.value
        return out

    def set_downsample(self, factor):
        """Downsampling factor (1, 2, 4, 8) of the following runs
        (pikgpu_set_downsample)."""
        self._check(self._lib.pikgpu_set_downsample(self._ctx, int(factor)))
        self._ds = int(factor)

    def set_output(self, outputs):
        """abi.OUTPUT_LINEAR | abi.OUTPUT_SRGB8 for the following runs."""
        self._check(self._lib.pikgpu_set_output(self._ctx, int(outputs)))

    def download_srgb8(self):
        """(ysize, xsize, 3) uint8 sRGB of the last run."""
        xs, ys = self._out_size()
        out = np.empty((ys, xs, 3), np.uint8)
        self._check(self._lib.pikgpu_download_srgb8(
            self._ctx, out.ctypes.data, xs * 3))
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

    def fwd_adj(self, qf, target=None, dl_dlinear=None, want_linear=False,
                want_dcoeffs=False):
        """Forward model + adjoint of the uploaded frame for a real-valued
        quant field (pikgpu_fwd_adj), host arrays. qf: (by, bx) float32 > 0;
        target or dl_dlinear: (3, ysize, xsize) float32. Returns a dict with
        dl_dqf (by, bx), and loss (with target), linear (3, ysize, xsize),
        dl_dcoeffs (3, by, bx*64) when asked for."""
        f = self._frame
        bx, by = f.xsize_blocks, f.ysize_blocks
        qf = np.ascontiguousarray(qf, np.float32)
        if qf.shape != (by, bx):
            raise ValueError("qf must be (%d, %d)" % (by, bx))
        src = target if target is not None else dl_dlinear
        if src is None:
            raise ValueError("need target or dl_dlinear")
        src = np.ascontiguousarray(src, np.float32)
        if src.shape != (3, f.ysize, f.xsize):
            raise ValueError("target / dl_dlinear must be (3, %d, %d)"
                             % (f.ysize, f.xsize))
        io = abi.AdjIoC()
        io.flags = 0
        io.qf = qf.ctypes.data
        io.qf_stride = bx
        planes = [src[c].ctypes.data for c in range(3)]
        if target is not None:
            io.target = (ctypes.c_void_p * 3)(*planes)
        else:
            io.dl_dlinear = (ctypes.c_void_p * 3)(*planes)
        io.in_stride = f.xsize
        out = {"dl_dqf": np.empty((by, bx), np.float32)}
        io.dl_dqf = out["dl_dqf"].ctypes.data
        io.dqf_stride = bx
        if want_linear:
            out["linear"] = np.empty((3, f.ysize, f.xsize), np.float32)
            io.linear = (ctypes.c_void_p * 3)(
                *[out["linear"][c].ctypes.data for c in range(3)])
            io.linear_stride = f.xsize
        if want_dcoeffs:
            out["dl_dcoeffs"] = np.empty((3, by, bx * 64), np.float32)
            io.dl_dcoeffs = (ctypes.c_void_p * 3)(
                *[out["dl_dcoeffs"][c].ctypes.data for c in range(3)])
            io.dcoeffs_stride = bx * 64
        loss = ctypes.c_double(0.0)
        if target is not None:
            io.loss = ctypes.addressof(loss)
        self._check(self._lib.pikgpu_fwd_adj(self._ctx, ctypes.byref(io)))
        if target is not None:
            out["loss"] = loss.value
        return out

    def fwd_adj_device(self, qf, dl_dqf, target=None, dl_dlinear=None,
                       linear=None, dl_dcoeffs=None, loss=None):
        """pikgpu_fwd_adj on device memory: every argument is a contiguous
        device array of the shapes fwd_adj uses (loss: one float64), e.g. a
        pik_amd.devmem.DeviceArray; anything with data_ptr() and, for the
        planes, first-axis indexing works. Asynchronous on the context
        stream (sync() waits)."""
        f = self._frame
        io = abi.AdjIoC()
        io.flags = abi.ADJ_DEVICE_POINTERS
        io.qf = qf.data_ptr()
        io.qf_stride = f.xsize_blocks
        src = target if target is not None else dl_dlinear
        planes = [src[c].data_ptr() for c in range(3)]
        if target is not None:
            io.target = (ctypes.c_void_p * 3)(*planes)
        else:
            io.dl_dlinear = (ctypes.c_void_p * 3)(*planes)
        io.in_stride = f.xsize
        io.dl_dqf = dl_dqf.data_ptr()
        io.dqf_stride = f.xsize_blocks
        if linear is not None:
            io.linear = (ctypes.c_void_p * 3)(
                *[linear[c].data_ptr() for c in range(3)])
            io.linear_stride = f.xsize
        if dl_dcoeffs is not None:
            io.dl_dcoeffs = (ctypes.c_void_p * 3)(
                *[dl_dcoeffs[c].data_ptr() for c in range(3)])
            io.dcoeffs_stride = f.xsize_blocks * 64
        if loss is not None:
            io.loss = loss.data_ptr()
        self._check(self._lib.pikgpu_fwd_adj(self._ctx, ctypes.byref(io)))

    def timings(self):
        names = (ctypes.c_char_p * 16)()
        ms = (ctypes.c_float * 16)()
        n = self._lib.pikgpu_last_timings(self._ctx, names, ms, 16)
        return {names[i].decode(): float(ms[i]) for i in range(n)}


def batch_phases():
    """Names of the batched reconstruction's phases, in launch order."""
    lib = abi.load()
    out = []
    for p in range(abi.BATCH_PHASES):
        out.append(lib.pikgpu_batch_phase_name(p, None).decode())
    return out


def batch_phase_can_persist():
    lib = abi.load()
    out = []
    for p in range(abi.BATCH_PHASES):
        c = ctypes.c_int()
        lib.pikgpu_batch_phase_name(p, ctypes.byref(c))
        out.append(bool(c.value))
    return out


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

    def schedule(self):
        """The batch's schedule: dict(mode, order, persist (per phase))."""
        sc = abi.ScheduleC()
        self._check(self._lib.pikgpu_batch_get_schedule(self._b,
                                                        ctypes.byref(sc)))
        return {"mode": sc.mode, "order": sc.order,
                "persist": [int(v) for v in sc.persist]}

    def set_schedule(self, mode=abi.BATCH_QUEUE, order=1, persist=None):
        """pikgpu_batch_set_schedule. persist: per-phase workgroups per CU
        of the phases that can loop (a list in phase order, or a dict
        phase name -> value; 0 = one workgroup per work item)."""
        sc = abi.ScheduleC()
        sc.mode, sc.order = int(mode), int(order)
        if isinstance(persist, dict):
            names = batch_phases()
            for k, v in persist.items():
                sc.persist[names.index(k)] = int(v)
        elif persist is not None:
            for i, v in enumerate(persist):
                sc.persist[i] = int(v)
        self._check(self._lib.pikgpu_batch_set_schedule(self._b,
                                                        ctypes.byref(sc)))

    def load_schedule(self, path=None):
        """Applies the schedule the autotuner chose (pik_amd/
        batch_schedule.json, or `path`); returns whether one was found."""
        import json
        import os
        path = path or os.path.join(os.path.dirname(os.path.abspath(
            __file__)), "batch_schedule.json")
        if not os.path.exists(path):
            return False
        with open(path) as f:
            d = json.load(f)
        self.set_schedule(d["mode"], d["order"], d.get("persist"))
        return True

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

    def add_pass(self, i, ac):
        """Adds an AC pass (3 planes of int16, the frame's AC layout) to
        frame i of the uploaded batch (pikgpu_batch_add_pass)."""
        f = self._frames[i]
        if len(ac) != 3:
            raise ValueError("a pass has three planes")
        want = (f.ysize_blocks, f.xsize_blocks * 64)
        planes = [np.ascontiguousarray(a, np.int16) for a in ac]
        for a in planes:
            if a.shape != want:
                raise ValueError("AC pass planes must be %s int16, got %s"
                                 % (want, a.shape))
        ptrs = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in planes])
        self._check(self._lib.pikgpu_batch_add_pass(
            self._b, int(i), ctypes.cast(ptrs, ctypes.c_void_p), want[1]))

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
