"""Synthetic inputs: seeded sRGB test images and a compact encoder that turns
them into decoder-state frames (pik_amd.frame.Frame).

The image generator follows the workload definitions of the benchmark
configurations: a sum of three smooth cosine gradients, Gaussian-blurred
uniform noise and solid rectangles with sharp edges (plus 2-px line art for
the large image).

The encoder is this repository's own: it is not the reference encoder (whose
heuristics cannot be linked here) and makes no attempt to match its
rate/distortion choices. It produces frames in exactly the reference's
decoder-state format -- variable-size transforms in the transposed-scaled
coefficient layout, per-block quant field, CfL maps, DC, dot placements --
with a strategy mix covering every AC strategy the decoder handles, so the
reconstruction path is exercised the same way a real bitstream would.
"""
import numpy as np

from . import abi
from .frame import DOT_DTYPE, Frame, div_ceil
from . import tables

# Opsin absorbance (pik/opsin_params.h:18-41): the forward transform is only
# needed to make test inputs.
_OPSIN_MIX = np.array([
    [0.29956550340058319, 0.63373087833825936, 0.077705617820981968],
    [0.22158691104574774, 0.68491388044116142, 0.096254234043612538],
    [0.20062661225219422, 0.070366199217588729, 0.5571760754215358],
]) / 255.0
_OPSIN_BIAS = np.array([0.26786006338144885, 0.24494032763907073,
                        0.14255999980363571]) / 255.0


def synthetic_srgb(xsize, ysize, seed, num_rects=20, line_art=False):
    """Seeded 8-bit sRGB image, shape (ysize, xsize, 3), uint8."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(ysize, dtype=np.float64) / ysize,
                         np.arange(xsize, dtype=np.float64) / xsize,
                         indexing="ij")
    img = np.empty((ysize, xsize, 3), np.float64)
    for c in range(3):
        acc = np.full((ysize, xsize), 128.0)
        for _ in range(3):
            fx, fy = rng.uniform(0.5, 4.0, size=2)
            ph = rng.uniform(0, 2 * np.pi)
            acc += 60.0 / 3.0 * np.cos(2 * np.pi * (fx * xx + fy * yy) + ph)
        noise = rng.uniform(-40.0, 40.0, size=(ysize, xsize))
        acc += gaussian_filter(noise, sigma=2.0, mode="reflect")
        img[:, :, c] = acc
    for _ in range(num_rects):
        w = int(rng.integers(max(2, xsize // 64), max(3, xsize // 6)))
        h = int(rng.integers(max(2, ysize // 64), max(3, ysize // 6)))
        x0 = int(rng.integers(0, max(1, xsize - w)))
        y0 = int(rng.integers(0, max(1, ysize - h)))
        img[y0:y0 + h, x0:x0 + w, :] = rng.uniform(0, 255, size=3)
    if line_art:
        for _ in range(max(8, (xsize * ysize) >> 16)):
            x0, x1 = rng.uniform(0, xsize, size=2)
            y0, y1 = rng.uniform(0, ysize, size=2)
            n = int(max(abs(x1 - x0), abs(y1 - y0))) + 1
            t = np.linspace(0.0, 1.0, n)
            px = (x0 + (x1 - x0) * t).astype(np.int64)
            py = (y0 + (y1 - y0) * t).astype(np.int64)
            col = rng.uniform(0, 255, size=3)
            for d in (0, 1):
                img[np.clip(py + d, 0, ysize - 1), np.clip(px, 0, xsize - 1)] = col
                img[np.clip(py, 0, ysize - 1), np.clip(px + d, 0, xsize - 1)] = col
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def srgb8_to_linear255(rgb):
    x = rgb.astype(np.float64) / 255.0
    lin = np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)
    return lin * 255.0


def linear_to_xyb(lin):
    """Linear RGB in [0,255] (H, W, 3) -> XYB planes (3, H, W) float64
    (LinearToXyb, pik/opsin_image.cc:34-47, with an exact cube root)."""
    mixed = np.einsum("ij,hwj->ihw", _OPSIN_MIX, lin) + _OPSIN_BIAS[:, None, None]
    mixed = np.cbrt(np.maximum(mixed, 0.0))
    r, g, b = mixed
    return np.stack([(r - g) * 0.5, (r + g) * 0.5, b])


# ---------------------------------------------------------------------------
# Transforms in the reference's transposed-scaled convention.
#
# Pixels P (N x N, [y][x]) and coefficients F (N x N, row-major in memory)
# relate by P = B^T F^T B with B[k][n] = w_N(k) cos(pi (2n+1) k / 2N): the
# coefficient row index is the horizontal frequency. w_N(0) = 1; for k > 0,
# w_8(k) = 1/cos(k pi/16) (the scaled Arai-Agui-Nakajima form the reference
# uses for DCT8) and w_N(k) = 2 cos(k pi/2N) for N = 4, 16, 32 (the
# Vashkevich-Petrovsky form). These follow from ComputeTransposedScaled[I]DCT
# (pik/dct.h:351-507) and its scale tables (pik/dct.h:25-130).

def _basis(n):
    k = np.arange(n)[:, None]
    x = np.arange(n)[None, :]
    if n == 8:
        w = np.where(k == 0, 1.0, 1.0 / np.cos(k * np.pi / 16.0))
    else:
        w = np.where(k == 0, 1.0, 2.0 * np.cos(k * np.pi / (2.0 * n)))
    return w * np.cos(np.pi * (2 * x + 1) * k / (2.0 * n))


_B = {n: _basis(n) for n in (4, 8, 16, 32)}
_BINV = {n: np.linalg.inv(_B[n]) for n in (4, 8, 16, 32)}


def fwd_dct(p):
    """Forward ts-DCT of pixel blocks (..., N, N) -> coefficients."""
    n = p.shape[-1]
    bi = _BINV[n]
    return np.einsum("yk,...xy,xl->...kl", bi, p, bi, optimize=True)


def inv_dct(f):
    n = f.shape[-1]
    b = _B[n]
    return np.einsum("cy,...rc,rx->...yx", b, f, b, optimize=True)


def _fwd_dct4x4(p):
    """DCT4X4 forward (pik/ac_strategy.cc:590-613) on (..., 8, 8)."""
    out = np.zeros(p.shape, np.float64)
    for y in range(2):
        for x in range(2):
            blk = fwd_dct(p[..., 4 * y:4 * y + 4, 4 * x:4 * x + 4])
            out[..., y::2, x::2] = blk
    b00, b01 = out[..., 0, 0].copy(), out[..., 0, 1].copy()
    b10, b11 = out[..., 1, 0].copy(), out[..., 1, 1].copy()
    out[..., 0, 0] = (b00 + b01 + b10 + b11) * 0.25
    out[..., 0, 1] = (b00 + b01 - b10 - b11) * 0.25
    out[..., 1, 0] = (b00 - b01 + b10 - b11) * 0.25
    out[..., 1, 1] = (b00 - b01 - b10 + b11) * 0.25
    return out


def _dct2_top(block, s):
    """One level of the DCT2X2 pyramid (pik/ac_strategy.cc:301-328)."""
    out = block.copy()
    h = s // 2
    c00 = block[..., 0:s:2, 0:s:2]
    c01 = block[..., 0:s:2, 1:s:2]
    c10 = block[..., 1:s:2, 0:s:2]
    c11 = block[..., 1:s:2, 1:s:2]
    out[..., :h, :h] = (c00 + c01 + c10 + c11) * 0.25
    out[..., :h, h:s] = (c00 + c01 - c10 - c11) * 0.25
    out[..., h:s, :h] = (c00 - c01 + c10 - c11) * 0.25
    out[..., h:s, h:s] = (c00 - c01 - c10 + c11) * 0.25
    return out


def _fwd_dct2x2(p):
    out = _dct2_top(p, 8)
    out = _dct2_top(out, 4)
    return _dct2_top(out, 2)


def _fwd_identity(p):
    """IDENTITY forward (pik/ac_strategy.cc:569-589) on (..., 8, 8)."""
    out = np.zeros(p.shape, np.float64)
    for y in range(2):
        for x in range(2):
            sub = p[..., 4 * y:4 * y + 4, 4 * x:4 * x + 4]
            dc = sub.mean(axis=(-1, -2))
            for iy in range(4):
                for ix in range(4):
                    if ix == 1 and iy == 1:
                        continue
                    out[..., y + iy * 2, x + ix * 2] = sub[..., iy, ix] - sub[..., 1, 1]
            out[..., y + 2, x + 2] = out[..., y, x]
            out[..., y, x] = dc
    b00, b01 = out[..., 0, 0].copy(), out[..., 0, 1].copy()
    b10, b11 = out[..., 1, 0].copy(), out[..., 1, 1].copy()
    out[..., 0, 0] = (b00 + b01 + b10 + b11) * 0.25
    out[..., 0, 1] = (b00 + b01 - b10 - b11) * 0.25
    out[..., 1, 0] = (b00 - b01 + b10 - b11) * 0.25
    out[..., 1, 1] = (b00 - b01 - b10 + b11) * 0.25
    return out


def forward_transform(raw, p):
    """Coefficients of pixel blocks p (..., S, S) for raw strategy `raw`, as
    an (..., S, S) array whose row-major flattening is the reference's
    coefficient order."""
    if raw in (abi.ACS_DCT, abi.ACS_DCT_NOHF, abi.ACS_DCT16X16,
               abi.ACS_DCT32X32):
        return fwd_dct(p)
    if raw in (abi.ACS_DCT4X4, abi.ACS_DCT4X4_NOHF):
        return _fwd_dct4x4(p)
    if raw == abi.ACS_DCT2X2:
        return _fwd_dct2x2(p)
    if raw == abi.ACS_IDENTITY:
        return _fwd_identity(p)
    if raw == abi.ACS_LINES:
        return _fwd_lines(p)
    raise ValueError("unsupported strategy %d" % raw)


# LINES (pik/ac_strategy.cc:380-568), row-major pixel / slot numbers: the
# diagonal lines (pixels, coefficient slots), the four mean groups and the
# slots of their means' Hadamard.
_LINES = [
    ([48, 56, 57], [48, 56, 57]), ([40, 49, 58], [40, 49, 58]),
    ([5, 14, 23], [5, 14, 23]), ([6, 7, 15], [6, 7, 15]),
    ([32, 41, 50, 59], [32, 41, 50, 59]), ([24, 33, 51, 60], [24, 33, 42, 51]),
    ([3, 12, 30, 39], [3, 12, 21, 30]), ([4, 13, 22, 31], [4, 13, 22, 31]),
    ([16, 25, 34, 43, 52, 61], [16, 25, 34, 43, 52, 61]),
    ([8, 17, 26, 44, 53, 62], [17, 26, 35, 44, 53, 62]),
    ([1, 10, 19, 37, 46, 55], [10, 19, 28, 37, 46, 55]),
    ([2, 11, 20, 29, 38, 47], [2, 11, 20, 29, 38, 47]),
    ([0, 9, 18, 27, 36, 45, 54, 63], [18, 27, 36, 45, 54, 60, 39, 63])]
_LINES_GROUPS = [
    [0, 1, 9, 2, 10, 18, 11, 19, 27, 20, 28, 36, 29, 37, 45, 38, 46, 54, 47,
     55, 63],
    [3, 4, 12, 5, 13, 21, 6, 14, 22, 30, 7, 15, 23, 31, 39],
    [24, 32, 40, 48, 56, 33, 41, 49, 57, 42, 50, 58, 51, 59, 60],
    [8, 16, 17, 25, 26, 34, 35, 43, 44, 52, 53, 61, 62]]
_LINES_SLOTS = [0, 8, 1, 9]


def _lines_dct(n, v):
    """The 1-D forward transforms of LINES (DCT3, DCT4, DCT6, DCT8,
    pik/ac_strategy.cc:66-257) in float64 on the last axis of v."""
    i = [v[..., k] for k in range(n)]
    if n == 3:
        return np.stack([(i[0] + i[1] + i[2]) / 3.0,
                         (i[0] - i[2]) * 0.57735029,
                         (i[0] + i[2]) / 3.0 - i[1] * (2 / 3.0)], -1)
    if n == 4:
        t0, t1, t2, t3 = i[0] + i[3], i[1] + i[2], i[0] - i[3], i[1] - i[2]
        t6, t7 = t2 - t3, t3 * 1.414213562373095048
        return np.stack([(t0 + t1) * 0.25, (t6 + t7) * (0.653281482438188264 / 2),
                         (t0 - t1) * 0.25, (t6 - t7) * (0.270598050073098492 / 2)],
                        -1)
    if n == 6:
        t0 = (i[1] - i[4]) * 0.23570227
        return np.stack([
            (i[0] + i[1] + i[2] + i[3] + i[4] + i[5]) / 6.0,
            (i[0] - i[5]) * 0.32197529 + t0 + (i[2] - i[3]) * 0.08627302,
            (i[0] - i[2] - i[3] + i[5]) * 0.28867514,
            (i[0] - i[1] - i[2] + i[3] + i[4] - i[5]) * 0.23570227,
            (i[0] + i[2] + i[3] + i[5]) / 6.0 - (i[1] + i[4]) / 3.0,
            (i[0] - i[5]) * 0.08627302 - t0 + (i[2] - i[3]) * 0.32197529], -1)
    c1, c2 = 0.707106781186548, 0.382683432365090
    c3, c4 = 1.30656296487638, 0.541196100146197
    t00, t01 = i[0] + i[7], i[0] - i[7]
    t02, t03 = i[3] + i[4], i[3] - i[4]
    t04, t05 = i[2] + i[5], i[2] - i[5]
    t06, t07 = i[1] + i[6], i[1] - i[6]
    t08, t09, t10, t11 = t00 + t02, t00 - t02, t06 + t04, t06 - t04
    t12, t13, t14 = t07 + t05, t01 + t07, t05 + t03
    t17, t18, t19 = c1 * (t11 + t09), c1 * t12, c2 * (t14 - t13)
    t20, t21 = t01 + t18, t01 - t18
    t22, t23 = c3 * t13 + t19, c4 * t14 + t19
    k = 2.8284271247461903
    return np.stack([(t08 + t10) * (0.353553390593273762 / k),
                     (t20 + t22) * (0.254897789552079584 / k),
                     (t09 + t17) * (0.270598050073098492 / k),
                     (t21 - t23) * (0.30067244346752264 / k),
                     (t08 - t10) * (0.353553390593273762 / k),
                     (t21 + t23) * (0.449988111568207852 / k),
                     (t09 - t17) * (0.653281482438188264 / k),
                     (t20 - t22) * (1.28145772387075309 / k)], -1)


def _fwd_lines(p):
    """Forward LINES of (..., 8, 8) pixel blocks (float64)."""
    flat = np.array(p, np.float64).reshape(p.shape[:-2] + (64,))
    out = np.zeros_like(flat)
    means = []
    for g in _LINES_GROUPS:
        m = flat[..., g].mean(axis=-1)
        flat[..., g] -= m[..., None]
        means.append(m)
    m0, m1, m2, m3 = means
    out[..., _LINES_SLOTS[0]] = (m0 + m1 + m2 + m3) * 0.25
    out[..., _LINES_SLOTS[1]] = (m0 - m1 + m2 - m3) * 0.25
    out[..., _LINES_SLOTS[2]] = (m0 + m1 - m2 - m3) * 0.25
    out[..., _LINES_SLOTS[3]] = (m0 - m1 - m2 + m3) * 0.25
    for pix, slots in _LINES:
        out[..., slots] = _lines_dct(len(pix), flat[..., pix])
    return out.reshape(p.shape)


def covered(raw):
    return 4 if raw == abi.ACS_DCT32X32 else (2 if raw == abi.ACS_DCT16X16 else 1)


# quant kinds (pik/quant_weights.h:25-33) by raw strategy
_KIND = {abi.ACS_DCT: 0, abi.ACS_IDENTITY: 1, abi.ACS_DCT2X2: 2,
         abi.ACS_DCT4X4: 3, abi.ACS_DCT16X16: 4, abi.ACS_DCT32X32: 5,
         abi.ACS_LINES: 6, abi.ACS_DCT_NOHF: 0, abi.ACS_DCT4X4_NOHF: 3}


def quant_kind(raw):
    return _KIND[raw]


# ---------------------------------------------------------------------------

def _hash01(a, b, seed):
    h = (np.asarray(a, np.uint64) * np.uint64(0x9E3779B97F4A7C15)
         ^ np.asarray(b, np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ np.uint64(seed * 0x165667B19E3779F9 & 0xFFFFFFFFFFFFFFFF))
    h ^= h >> np.uint64(29)
    h *= np.uint64(0xBF58476D1CE4E5B9)
    h ^= h >> np.uint64(32)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(1 << 24)


def choose_strategies(xyb_pad, seed, lines=0.0):
    """Per-block raw strategy map (by, bx) uint8 from local activity, with a
    deterministic mix that covers every strategy the reference encoder
    emits; `lines` is the fraction of the single blocks coded with LINES,
    which the reference decodes but its encoder never chooses
    (pik/ac_strategy.cc:18)."""
    _, h, w = xyb_pad.shape
    by, bx = h // 8, w // 8
    y = xyb_pad[1].reshape(by, 8, bx, 8)
    act = y.std(axis=(1, 3))  # per-block activity
    raw = np.full((by, bx), 255, np.int32)
    r = _hash01(np.arange(by)[:, None], np.arange(bx)[None, :], seed)
    # 32x32 on smooth aligned regions
    for s, code, thr, prob in ((4, abi.ACS_DCT32X32, 0.012, 0.85),
                               (2, abi.ACS_DCT16X16, 0.025, 0.85)):
        for y0 in range(0, by - s + 1, s):
            for x0 in range(0, bx - s + 1, s):
                if (raw[y0:y0 + s, x0:x0 + s] != 255).any():
                    continue
                if act[y0:y0 + s, x0:x0 + s].max() < thr and r[y0, x0] < prob:
                    raw[y0:y0 + s, x0:x0 + s] = code + 1000  # marker
    free = raw == 255
    single = np.full((by, bx), abi.ACS_DCT, np.int32)
    hi = act > 0.05
    single[hi & (r < 0.35)] = abi.ACS_DCT4X4
    single[hi & (r >= 0.35) & (r < 0.45)] = abi.ACS_IDENTITY
    single[hi & (r >= 0.45) & (r < 0.55)] = abi.ACS_DCT2X2
    single[(~hi) & (r < 0.05)] = abi.ACS_DCT_NOHF
    single[(~hi) & (r >= 0.05) & (r < 0.09)] = abi.ACS_DCT4X4_NOHF
    single[(~hi) & (r >= 0.09) & (r < 0.13)] = abi.ACS_DCT2X2
    single[(~hi) & (r >= 0.13) & (r < 0.16)] = abi.ACS_IDENTITY
    single[(~hi) & (r >= 0.16) & (r < 0.22)] = abi.ACS_DCT4X4
    if lines > 0:
        r2 = _hash01(np.arange(by)[:, None], np.arange(bx)[None, :], seed + 77)
        single[r2 < lines] = abi.ACS_LINES
    raw = np.where(free, single, raw)
    out = np.zeros((by, bx), np.uint8)
    for code in (abi.ACS_DCT32X32, abi.ACS_DCT16X16):
        s = covered(code)
        yy, xx = np.nonzero(raw == code + 1000)
        # groups are s-aligned, so the index inside the transform is local
        out[yy, xx] = (code << 4) | ((yy % s) * s + (xx % s))
    single_mask = raw < 1000
    out[single_mask] = (raw[single_mask] << 4).astype(np.uint8)
    return out


# Quantizer::SetQuantField's target for the median quant value
# (pik/quantizer.h:155).
QUANT_FIELD_TARGET = 3.80987740592518214386


def set_quant_field(qf_float, quant_dc_float, global_scale=None):
    """Integer quantizer state from a float quant field the way
    Quantizer::SetQuantField (pik/quantizer.h:142-199) derives it: the
    global scale from the field's median minus its median absolute deviation
    (capped by quant_dc * 4096 and to [1, 32768]), then every value times
    65536 / global_scale rounded and clamped to [1, 256]. An explicit
    `global_scale` replaces the derived one. Returns (global_scale,
    quant_dc, int32 field)."""
    q = np.asarray(qf_float, np.float32)
    if global_scale is None:
        med = np.float32(np.median(q))
        mad = np.float32(np.median(np.abs(q - med)))
        gs = int(np.float32(65536.0) * (med - mad)
                 / np.float32(QUANT_FIELD_TARGET))
        gs = min(gs, int(quant_dc_float * 4096))
        global_scale = max(1, min(gs, 1 << 15))
    inv_gs = np.float32(65536.0 / global_scale)
    quant_dc = int(np.clip(int(np.float32(quant_dc_float) * inv_gs + 0.5),
                           1, 256))
    field = np.clip((q * inv_gs + np.float32(0.5)).astype(np.int64), 1,
                    256).astype(np.int32)
    return int(global_scale), quant_dc, field


def max_over_transforms(field, ac_strategy):
    """A float quant field with every multi-block transform's blocks set to
    their maximum (AdjustQuantField, pik/adaptive_quantization.cc:554-578)."""
    out = np.array(field, np.float32, copy=True)
    raw = ac_strategy >> 4
    first = (ac_strategy & 15) == 0
    for code in (abi.ACS_DCT16X16, abi.ACS_DCT32X32):
        s = covered(code)
        for y0, x0 in zip(*np.nonzero(first & (raw == code))):
            out[y0:y0 + s, x0:x0 + s] = out[y0:y0 + s, x0:x0 + s].max()
    return out


def encode_frame(rgb, distance=1.0, seed=0, num_dots=0,
                 lut_ids="zero", predict=True, grayscale=False,
                 global_scale=None, lines=0.0):
    """Encodes an (H, W, 3) uint8 sRGB image into a Frame. The quantizer
    scalars come from a float quant field as the reference encoder derives
    them (set_quant_field), so global_scale is in general no power of two;
    an explicit `global_scale` overrides it. With
    `grayscale` the image is first made gray (channel mean) and the frame
    is a kGrayscaleOpt frame: only the Y DC is coded, the X and B DC planes
    hold zeros as Image3SDecompress leaves them (pik/compressed_dc.cc:
    107-121), and the decoder rebuilds them from Y."""
    rgb = np.asarray(rgb)
    if grayscale:
        g = np.rint(rgb.astype(np.float64).mean(axis=2)).astype(np.uint8)
        rgb = np.repeat(g[:, :, None], 3, axis=2)
    ysize, xsize = rgb.shape[:2]
    bx, by = div_ceil(xsize, 8), div_ceil(ysize, 8)
    tx, ty = div_ceil(xsize, 64), div_ceil(ysize, 64)
    xyb = linear_to_xyb(srgb8_to_linear255(rgb))
    # PadImageToMultiple (pik/image.cc:45-80): replicate last column/row.
    xyb = np.pad(xyb, ((0, 0), (0, by * 8 - ysize), (0, bx * 8 - xsize)),
                 mode="edge")
    table, offsets = tables.default_tables()
    acs = choose_strategies(xyb, seed, lines)
    raw = (acs >> 4).astype(np.int32)
    first = (acs & 15) == 0

    # Quant field: finer where the block is active; coarser with distance.
    yb = xyb[1].reshape(by, 8, bx, 8)
    act = yb.std(axis=(1, 3))
    rng = np.random.default_rng(seed + 17)
    qf_float = ((5.0 + 60.0 * act + rng.uniform(0, 4, act.shape))
                / (8.0 * max(distance, 0.05))).astype(np.float32)
    # AdjustQuantField semantics: a multi-block transform uses its maximum.
    for code in (abi.ACS_DCT16X16, abi.ACS_DCT32X32):
        s = covered(code)
        ys, xs = np.nonzero(first & (raw == code))
        for y0, x0 in zip(ys, xs):
            qf_float[y0:y0 + s, x0:x0 + s] = \
                qf_float[y0:y0 + s, x0:x0 + s].max()
    global_scale, quant_dc, qf = set_quant_field(
        qf_float, 6.0 / max(distance, 0.05), global_scale)
    inv_gs = np.float32(65536.0 / global_scale)
    inv_quant_dc = np.float32(inv_gs / np.float32(quant_dc))

    # Coefficients per strategy, in float64.
    coef = np.zeros((3, by, bx * 64), np.float64)
    for code in (abi.ACS_DCT, abi.ACS_IDENTITY, abi.ACS_DCT2X2,
                 abi.ACS_DCT4X4, abi.ACS_DCT16X16, abi.ACS_DCT32X32,
                 abi.ACS_DCT_NOHF, abi.ACS_DCT4X4_NOHF, abi.ACS_LINES):
        s = covered(code)
        ys, xs = np.nonzero(first & (raw == code))
        if len(ys) == 0:
            continue
        S = 8 * s
        iy = (ys[:, None] * 8 + np.arange(S)[None, :])
        ix = (xs[:, None] * 8 + np.arange(S)[None, :])
        pix = xyb[:, iy[:, :, None], ix[:, None, :]]  # (3, n, S, S)
        f = forward_transform(code, pix).reshape(3, len(ys), s, 64 * s)
        for r in range(s):
            rows = ys + r
            cols = xs[:, None] * 64 + np.arange(64 * s)[None, :]
            coef[:, rows[:, None], cols] = f[:, :, r, :]

    # CfL factors per 64x64 tile from least squares on the AC coefficients.
    ytox = np.full((ty, tx), 128, np.int32)
    ytob = np.full((ty, tx), 120, np.int32)
    cy = coef[1].reshape(by, bx, 64)
    for t_y in range(ty):
        for t_x in range(tx):
            sl = (slice(8 * t_y, 8 * t_y + 8), slice(8 * t_x, 8 * t_x + 8))
            yy = cy[sl][..., 1:].ravel()
            den = float((yy * yy).sum())
            if den > 1e-9:
                xr = coef[0].reshape(by, bx, 64)[sl][..., 1:].ravel()
                br = coef[2].reshape(by, bx, 64)[sl][..., 1:].ravel()
                ytox[t_y, t_x] = int(np.clip(round(128 + 256 * (xr @ yy) / den), 0, 255))
                ytob[t_y, t_x] = int(np.clip(round(128 * (br @ yy) / den), 0, 255))

    # Quantize AC: step = M[k] * inv_gs / qf; X/B residual after CfL.
    kind = np.vectorize(quant_kind)(raw)
    blk_index = (acs & 15).astype(np.int64)
    off = offsets.reshape(-1, 7, 3)[0]  # dequant table 0
    k = np.arange(64)
    ac = np.zeros((3, by, bx * 64), np.int16)
    deq_y = None
    tile_y = np.arange(by) // 8
    tile_x = np.arange(bx) // 8
    fx = ((ytox[tile_y[:, None], tile_x[None, :]] - 128) / 256.0)
    fb = (ytob[tile_y[:, None], tile_x[None, :]] / 128.0)
    scale = inv_gs / qf  # (by, bx)
    for c in (1, 0, 2):
        mats = table[off[kind, c][..., None] + blk_index[..., None] * 64 + k]
        step = (mats * scale[..., None]).astype(np.float64)  # (by, bx, 64)
        v = coef[c].reshape(by, bx, 64)
        if c != 1:
            mul = fx if c == 0 else fb
            v = v - mul[..., None] * deq_y
        q = np.clip(np.rint(v / step), -32767, 32767)
        q[..., 0] = 0  # DC travels separately
        ac[c] = q.reshape(by, bx * 64).astype(np.int16)
        if c == 1:
            qa = q.copy()
            adj = np.where(np.abs(qa) >= 1.5, qa - 0.145 / np.where(qa == 0, 1, qa),
                           np.where(np.abs(qa) >= 0.5,
                                    np.sign(qa) * (1 - 0.07005449891748593), 0.0))
            deq_y = adj * step

    # DC: per-8x8-block means, quantized with mul_dc = M_dct8[0] * inv_quant_dc.
    dcimg = xyb.reshape(3, by, 8, bx, 8).mean(axis=(2, 4))
    mul_dc = np.array([table[off[0, c]] for c in range(3)], np.float32) * inv_quant_dc
    dc = np.zeros((3, by, bx), np.int16)
    qy = np.rint(dcimg[1] / mul_dc[1])
    dc[1] = np.clip(qy, -32767, 32767)
    ydec = dc[1].astype(np.float64) * mul_dc[1]
    ytox_dc = int(np.clip(round(128 + 256 * float((dcimg[0] * ydec).sum())
                                / max(float((ydec * ydec).sum()), 1e-9)), 0, 255))
    ytob_dc = int(np.clip(round(128 * float((dcimg[2] * ydec).sum())
                                / max(float((ydec * ydec).sum()), 1e-9)), 0, 255))
    for c, f in ((0, (ytox_dc - 128) / 256.0), (2, ytob_dc / 128.0)):
        dc[c] = np.clip(np.rint((dcimg[c] - f * ydec) / mul_dc[c]), -32767, 32767)
    if grayscale:
        dc[0] = 0
        dc[2] = 0

    if lut_ids == "zero":
        lut = np.zeros((by, bx), np.uint8)
    else:
        lut = (_hash01(np.arange(by)[:, None], np.arange(bx)[None, :],
                       seed + 5) < 0.3).astype(np.uint8)

    dots = make_dots(xsize, ysize, num_dots, seed + 99)
    return Frame(
        xsize=xsize, ysize=ysize, ac=ac, dc=dc, quant_field=qf,
        ac_strategy=acs, sigma_lut_ids=lut,
        dequant_cf=np.zeros((ty, tx), np.uint8), ytox_map=ytox, ytob_map=ytob,
        dequant_map=np.zeros((16, 256), np.uint8), dequant_table=table,
        dequant_offsets=offsets, dots=dots, global_scale=global_scale,
        quant_dc=quant_dc, ytox_dc=ytox_dc, ytob_dc=ytob_dc,
        predict_lf=int(predict), predict_hf=int(predict),
        grayscale=int(grayscale)).contiguous()


def make_dots(xsize, ysize, n, seed):
    """n dot placements: small int8 patches, a third of them with an affine
    (scaled/rotated) splat, all inside the image."""
    dots = np.zeros(n, DOT_DTYPE)
    rng = np.random.default_rng(seed)
    for i in range(n):
        xs, ys = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        d = dots[i]
        d["xsize"], d["ysize"] = xs, ys
        d["x"] = int(rng.integers(0, max(1, xsize - xs - 8)))
        d["y"] = int(rng.integers(0, max(1, ysize - ys - 8)))
        pix = np.zeros((3, 64), np.int8)
        pix[:, :xs * ys] = rng.integers(-12, 13, size=(3, xs * ys))
        d["pixels"] = pix
        if i % 3 == 2:
            d["transform"] = 1
            d["dx"] = int(rng.integers(2, 7))
            d["dy"] = int(rng.integers(-3, 4))
            d["width"] = int(rng.integers(2, 9))
    return dots


def make_frame(xsize, ysize, seed=0, distance=1.0, num_rects=20,
               line_art=False, **kw):
    rgb = synthetic_srgb(xsize, ysize, seed, num_rects=num_rects,
                         line_art=line_art)
    return encode_frame(rgb, distance=distance, seed=seed, **kw)


def make_opsin_frame(xsize, ysize, seed=0, distance=1.0, num_rects=20,
                     line_art=False, predict=True, lut_ids="zero", lines=0.0):
    """A frame's side information together with the XYB image it was made
    from, padded to whole blocks by edge replication as PadImageToMultiple
    does (pik/image.cc:45-80): (frame, opsin (3, by * 8, bx * 8) float32).
    The encoder-side quantization takes both."""
    rgb = synthetic_srgb(xsize, ysize, seed, num_rects=num_rects,
                         line_art=line_art)
    frame = encode_frame(rgb, distance=distance, seed=seed, predict=predict,
                         lut_ids=lut_ids, lines=lines)
    xyb = linear_to_xyb(srgb8_to_linear255(rgb))
    bx, by = div_ceil(xsize, 8), div_ceil(ysize, 8)
    xyb = np.pad(xyb, ((0, 0), (0, by * 8 - ysize), (0, bx * 8 - xsize)),
                 mode="edge")
    return frame, np.ascontiguousarray(xyb, np.float32)


# The encoder's pre-sharpening before quantization when Gaborish is on
# (GaborishInverse, pik/gaborish.cc:92-119; applied at pik/pik_frame.cc:
# 527-529): a 5x5 symmetric kernel, normalised to unit sum, with clamped
# borders. Quadrant weights [|dy|][|dx|] before scaling by 0.92718927264540152
# (the centre stays 1).
_GAB_INV = (-0.092359145662814029, -0.039253623634014627, 0.016176494530216929,
            0.00083458437774987476, 0.004512465323949319)
_GAB_INV_MUL = 0.92718927264540152


def gaborish_inverse(xyb):
    """GaborishInverse of a (3, H, W) image (float32 weights, float64
    accumulation)."""
    g = [np.float32(_GAB_INV_MUL * v) for v in _GAB_INV]
    quad = np.array([[1.0, g[0], g[2]],
                     [g[0], g[1], g[3]],
                     [g[2], g[3], g[4]]], np.float32)
    w = np.zeros((5, 5), np.float64)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            w[dy + 2, dx + 2] = quad[abs(dy), abs(dx)]
    w = (w * (1.0 / w.sum())).astype(np.float32).astype(np.float64)
    img = np.asarray(xyb, np.float64)
    _, h, wd = img.shape
    pad = np.pad(img, ((0, 0), (2, 2), (2, 2)), mode="edge")
    out = np.zeros_like(img)
    for dy in range(5):
        for dx in range(5):
            out += w[dy, dx] * pad[:, dy:dy + h, dx:dx + wd]
    return out.astype(np.float32)


def make_search_inputs(xsize, ysize, seed=0, num_rects=20, line_art=False,
                       target=1.0):
    """Inputs of a quant-field search at Butteraugli distance `target`: the
    frame's side information (no LF / HF predictions, no dots; AC strategy,
    colour correlation and sigma LUTs from this generator), the padded XYB
    image the encoder quantizes (pre-sharpened by GaborishInverse as the
    reference encoder does with Gaborish on), the unpadded XYB original and
    its linear RGB (3, H, W). Returns (frame, opsin, opsin_orig, linear)."""
    rgb = synthetic_srgb(xsize, ysize, seed, num_rects=num_rects,
                         line_art=line_art)
    frame = encode_frame(rgb, distance=target, seed=seed, predict=False)
    lin = srgb8_to_linear255(rgb)
    xyb = linear_to_xyb(lin).astype(np.float32)
    linear = np.ascontiguousarray(lin.transpose(2, 0, 1), np.float32)
    bx, by = div_ceil(xsize, 8), div_ceil(ysize, 8)
    padded = np.pad(xyb, ((0, 0), (0, by * 8 - ysize), (0, bx * 8 - xsize)),
                    mode="edge")
    opsin = gaborish_inverse(padded) if frame.gaborish else padded
    return (frame, np.ascontiguousarray(opsin, np.float32),
            np.ascontiguousarray(xyb, np.float32),
            np.ascontiguousarray(linear, np.float32))


def make_extra_passes(frame, num_passes, seed=0, density=0.02, amplitude=1):
    """AC refinement passes for `frame` (multi-pass accumulation): each a
    (3, by, bx * 64) int16 array of sparse small values on the frame's
    coefficient layout, zero at every block's DC position (the AC image
    never carries the DC)."""
    rng = np.random.default_rng(seed)
    by, bx = frame.ysize_blocks, frame.xsize_blocks
    out = []
    for _ in range(num_passes):
        v = rng.integers(-amplitude, amplitude + 1, (3, by, bx * 64))
        v[rng.random(v.shape) > density] = 0
        v[:, :, ::64] = 0
        out.append(v.astype(np.int16))
    return out


def split_passes(frame, num_passes, seed=0):
    """Splits the frame's AC over num_passes passes, every coefficient in
    exactly one of them: returns (frame with the first pass's AC, list of
    the later passes). Each position then accumulates one dequantised value
    and zeros (x + 0 = x), and the colour correlation only reorders a
    commutative addition, so the multi-pass decode equals the single-pass
    one bit for bit."""
    rng = np.random.default_rng(seed)
    ac = np.asarray(frame.ac)
    which = rng.integers(0, num_passes, ac.shape)
    first = frame.copy()
    first.ac = np.where(which == 0, ac, 0).astype(np.int16)
    extra = [np.where(which == p, ac, 0).astype(np.int16)
             for p in range(1, num_passes)]
    return first, extra
