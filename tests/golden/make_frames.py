"""Writes tests/golden/frames.npz: whole frames (FrameHeader to the last
pass's groups) written by the reference encoder's pieces in PixelsToPikPass'
order (pik/pik_frame.cc:555-681: pass header, quantization tables,
quantizer, colour maps, DC groups, dot dictionary, dequant control field and
map, then per pass the group size table and the groups with their group
headers) from pik_amd.synth frames, with what the reference decoder's pieces
in PikPassToPixels' order make of them (ReadPassHeader,
DequantMatrices::Decode, Quantizer::Decode, DecodeColorMap x 2, the DC
groups, BlockDictionary::Decode, DecodeDequantControlField(Map), the group
size tables, ReadGroupHeader and DecodeFromBitstream's reading per group and
pass): header fields, quantized DC, side information, colour maps, dequant
control field and map, dots, every pass's AC, noise parameters and the
frame's length.

The reference code is built under oracle/_ref (make -C oracle ref); run
from the repository root:

    python3 tests/golden/make_frames.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pik_amd import abi, synth  # noqa: E402

# name: recipe. kw: make_frame arguments; mods: frame fields set afterwards;
# passes: number of further AC passes (make_extra_passes); noise: per-group
# noise parameters; variant: ref.VAR_* bits; table1: dequant map entries
# that the frame uses alternate between table sets 0 and 1;
# transform_dots_only: keeps the transformed dots only (the reference
# encoder writes transformed dots only in frames without plain ones).
CASES = {
    "tiny": dict(w=13, h=7, seed=2),
    "dots": dict(w=200, h=136, seed=4, kw=dict(num_dots=5),
                 variant=ref.VAR_STATIC_DOTS | ref.VAR_FRAME_INFO),
    "tdots": dict(w=264, h=264, seed=13, kw=dict(num_dots=12),
                  transform_dots_only=True, variant=ref.VAR_STATIC_DOTS),
    "passes": dict(w=520, h=264, seed=3, kw=dict(lines=0.2), passes=2,
                   noise=True,
                   variant=(ref.VAR_GROUP_HEADERS | ref.VAR_DOWNSAMPLING |
                            ref.VAR_EXTENSIONS)),
    "gray": dict(w=264, h=200, seed=11, kw=dict(grayscale=True),
                 variant=ref.VAR_FAST_AC),
    "tables": dict(w=300, h=200, seed=8, variant=ref.VAR_COPY_TABLE,
                   table1=True),
    "plain": dict(w=1032, h=600, seed=9,
                  mods=dict(gaborish=0, predict_lf=0, predict_hf=0,
                            adaptive_reconstruction=0),
                  variant=ref.VAR_GROUP_HEADERS),
    "sharp": dict(w=136, h=520, seed=12,
                  mods=dict(epf_use_sharpened=1, gaborish=4)),
}


def make(name):
    r = CASES[name]
    fr = synth.make_frame(r["w"], r["h"], seed=r["seed"], **r.get("kw", {}))
    for k, v in r.get("mods", {}).items():
        setattr(fr, k, v)
    if fr.grayscale:
        fr.flags |= abi.FLAG_GRAYSCALE_OPT
    if r.get("transform_dots_only"):
        fr.dots = fr.dots[fr.dots["transform"] == 1].copy()
        assert len(fr.dots) > 0
    if r.get("table1"):
        # alternate the used (control value, quant) entries between sets
        used = sorted({(int(fr.dequant_cf[y // 8, x // 8]),
                        int(fr.quant_field[y, x]) - 1)
                       for y in range(fr.ysize_blocks)
                       for x in range(fr.xsize_blocks)})
        for i, (cf, q) in enumerate(used):
            fr.dequant_map[cf, q] = i % 2
    extra = synth.make_extra_passes(fr, r.get("passes", 0), seed=r["seed"])
    ng = (-(-fr.xsize // 512)) * (-(-fr.ysize // 512))
    noise = None
    if r.get("noise"):
        rng = np.random.default_rng(r["seed"])
        noise = (rng.integers(-30000, 30000, (1 + len(extra), ng, 3)) /
                 1000.0).astype(np.float32)
    return fr, extra, noise, r.get("variant", 0)


def main():
    out = {}
    for name, r in CASES.items():
        fr, extra, noise, variant = make(name)
        data = ref.encode_frame(fr, extra, noise, variant)
        got = ref.decode_frame(data, fr.xsize, fr.ysize)
        # Lossless on what the format carries (the AC is not: the
        # low-frequency corner of a multi-block transform is not coded).
        assert np.array_equal(got["dc"], fr.dc), name
        assert np.array_equal(got["quant_field"], fr.quant_field), name
        assert got["consumed"] == len(data), name
        out[name + "/recipe"] = np.array(json.dumps(r))
        out[name + "/data"] = np.frombuffer(data, np.uint8)
        out[name + "/header"] = np.array(
            [got[k] for k in ref.FRAME_HEADER_FIELDS], np.int64)
        for k in ("dc", "ac_strategy", "quant_field", "sigma_lut_ids",
                  "ytox_map", "ytob_map", "dequant_cf", "dequant_map",
                  "dots", "noise"):
            out[name + "/" + k] = got[k]
        out[name + "/passes"] = np.stack(got["passes"])
        out[name + "/consumed"] = np.array(got["consumed"])
        print(name, len(data), flush=True)
    np.savez_compressed(os.path.join(HERE, "frames.npz"), **out)


if __name__ == "__main__":
    main()
