"""Frame: the decoder state after entropy decoding, as numpy arrays.

Field meanings follow include/pikgpu.h (pikgpu_frame), which mirrors the
reference's FrameDecCache / GroupDecCache contents after DecodeDCGroups and
DecodeFromBitstream (pik/compressed_image_fwd.h:104-229).
"""
import ctypes
import dataclasses

import numpy as np

from . import abi

DOT_DTYPE = np.dtype([
    ("x", "<u4"), ("y", "<u4"), ("xsize", "<u4"), ("ysize", "<u4"),
    ("transform", "<i4"), ("dx", "<i4"), ("dy", "<i4"), ("width", "<i4"),
    ("pixels", "i1", (3, 64)),
])
assert DOT_DTYPE.itemsize == ctypes.sizeof(abi.Dot)


def div_ceil(a, b):
    return -(-a // b)


@dataclasses.dataclass
class Frame:
    xsize: int
    ysize: int
    ac: np.ndarray            # (3, by, bx*64) int16
    dc: np.ndarray            # (3, by, bx) int16
    quant_field: np.ndarray   # (by, bx) int32, 1..256
    ac_strategy: np.ndarray   # (by, bx) uint8, (type<<4)|idx
    sigma_lut_ids: np.ndarray  # (by, bx) uint8
    dequant_cf: np.ndarray    # (ty, tx) uint8
    ytox_map: np.ndarray      # (ty, tx) int32
    ytob_map: np.ndarray      # (ty, tx) int32
    dequant_map: np.ndarray   # (16, 256) uint8
    dequant_table: np.ndarray  # float32
    dequant_offsets: np.ndarray  # uint32
    dots: np.ndarray          # structured DOT_DTYPE
    global_scale: int
    quant_dc: int
    ytox_dc: int = 128
    ytob_dc: int = 120
    gaborish: int = abi.GABORISH_750
    predict_lf: int = 1
    predict_hf: int = 1
    adaptive_reconstruction: int = 1
    epf_enable_adaptive: int = 1
    epf_use_sharpened: int = 0
    grayscale: int = 0
    flags: int = 0            # FrameHeader::flags (abi.FLAG_*)

    @property
    def xsize_blocks(self):
        return div_ceil(self.xsize, 8)

    @property
    def ysize_blocks(self):
        return div_ceil(self.ysize, 8)

    @property
    def pixels(self):
        return self.xsize * self.ysize

    def copy(self):
        """Deep copy (arrays copied)."""
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.copy() if isinstance(v, np.ndarray) else v
        return Frame(**kw)

    def contiguous(self):
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray):
                setattr(self, f.name, np.ascontiguousarray(v))
        return self

    def rows(self):
        """Makes every array row-addressable for the C ABI: unit stride
        along a row, rows at a positive stride that is a whole number of
        elements. Views into padded buffers (the reference's Image layout,
        rows of BytesPerRow bytes, pik/image.h:47-68) keep their stride;
        anything else is copied to a contiguous array."""
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if not isinstance(v, np.ndarray):
                continue
            ok = v.ndim >= 1 and (v.shape[-1] <= 1 or
                                  v.strides[-1] == v.itemsize)
            if ok and v.ndim >= 2:
                rs = v.strides[-2]
                ok = (rs > 0 and rs % v.itemsize == 0 and
                      rs >= v.shape[-1] * v.itemsize)
            if ok and v.ndim == 1:
                ok = v.flags.c_contiguous
            if not ok:
                setattr(self, f.name, np.ascontiguousarray(v))
        return self

    @staticmethod
    def _stride(a):
        """Row stride of a row-addressable array, in elements."""
        return a.strides[-2] // a.itemsize if a.shape[-2] > 1 else a.shape[-1]

    def to_c(self):
        """Returns (FrameC, keepalive) with pointers into this frame; row
        strides are taken from the arrays (rows())."""
        self.rows()
        f = abi.FrameC()
        f.xsize, f.ysize = self.xsize, self.ysize
        for name in ("gaborish", "predict_lf", "predict_hf",
                     "adaptive_reconstruction", "epf_enable_adaptive",
                     "epf_use_sharpened", "global_scale", "quant_dc",
                     "ytox_dc", "ytob_dc", "grayscale", "flags"):
            setattr(f, name, int(getattr(self, name)))
        for c in range(3):
            f.ac[c] = self.ac[c].ctypes.data
            f.dc[c] = self.dc[c].ctypes.data
        st = self._stride
        f.ac_stride = st(self.ac)
        f.dc_stride = st(self.dc)
        f.quant_field = self.quant_field.ctypes.data
        f.quant_field_stride = st(self.quant_field)
        f.ac_strategy = self.ac_strategy.ctypes.data
        f.ac_strategy_stride = st(self.ac_strategy)
        f.sigma_lut_ids = self.sigma_lut_ids.ctypes.data
        f.sigma_lut_stride = st(self.sigma_lut_ids)
        f.dequant_cf = self.dequant_cf.ctypes.data
        f.dequant_cf_stride = st(self.dequant_cf)
        f.ytox_map = self.ytox_map.ctypes.data
        f.ytox_stride = st(self.ytox_map)
        f.ytob_map = self.ytob_map.ctypes.data
        f.ytob_stride = st(self.ytob_map)
        if not self.dequant_map.flags.c_contiguous:
            self.dequant_map = np.ascontiguousarray(self.dequant_map)
        f.dequant_map = self.dequant_map.ctypes.data
        f.dequant_table = self.dequant_table.ctypes.data
        f.dequant_table_size = self.dequant_table.size
        f.dequant_offsets = self.dequant_offsets.ctypes.data
        f.num_dequant_offsets = self.dequant_offsets.size
        if len(self.dots):
            f.dots = self.dots.ctypes.data
        f.num_dots = len(self.dots)
        return f, self

    def save(self, path):
        d = {}
        for fld in dataclasses.fields(self):
            v = getattr(self, fld.name)
            d[fld.name] = v if isinstance(v, np.ndarray) else np.array(v)
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        kw = {}
        for fld in dataclasses.fields(cls):
            if fld.name not in z.files:
                continue  # written before the field existed: its default
            v = z[fld.name]
            kw[fld.name] = v if v.ndim else v.item()
        return cls(**kw).contiguous()
