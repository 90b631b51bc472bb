"""Writes tests/golden/dc_groups.npz: DC groups sections written by the
reference encoder's pieces in EncodeDCGroups' order (pik/compressed_dc.cc:
233-273, 352-433: per 256x256-block group ShrinkDC + EncodeImageData, then
the AC strategy, quant field and AR parameter tokens under one set of
histograms; a SizeCoderT size table in front) from pik_amd.synth frames,
with what the reference decoder's pieces in DecodeDCGroups' order make of
them (DecodeImage, ExpandDC, DecodeAcStrategy, DecodeQuantField,
DecodeARParameters): quantized DC, strategy bytes, quant field, sigma LUT
ids and the section length.

Most cases use the reference's portable scalar build. `wide_dc` (DC values
up to +-10000) uses its default AVX2 build: there the DC predictor's costs
reach 32768 and the two builds pick different predictors (the AVX2 build
compares costs as unsigned 16-bit values, the scalar one as signed), so
they write different streams; pikgpu decodes as the default build does.
The reference code is built under oracle/_ref (make -C oracle ref); run
from the repository root:

    python3 tests/golden/make_dc_groups.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pik_amd import synth  # noqa: E402

# name: (w, h, seed, make_frame kwargs, options)
CASES = {
    "tiny": (13, 7, 2, {}, {}),
    "small_noise": (200, 136, 4, {}, {}),
    "gray": (264, 200, 11, dict(grayscale=True), {}),
    "lines_groups": (600, 520, 3, dict(lines=0.2), {}),
    "groups_fast": (1032, 600, 9, {}, {}),
    "lines_lut": (600, 520, 3, dict(lines=0.2, lut_ids="rand"), {}),
    "two_groups": (2100, 300, 12, dict(distance=2.0), {}),
    "narrow": (8, 520, 13, {}, {}),
    "three_wide": (24, 24, 14, {}, {}),
    "wide_dc": (300, 200, 5, {}, dict(dc_amplitude=10000, kind="avx2")),
}


def frame_of(w, h, seed, kw, opts):
    fr = synth.make_frame(w, h, seed=seed, **kw)
    amp = opts.get("dc_amplitude")
    if amp:
        rng = np.random.default_rng(seed)
        fr.dc = rng.integers(-amp, amp + 1, size=fr.dc.shape).astype(np.int16)
    return fr


def main():
    out = {}
    for name, (w, h, seed, kw, opts) in CASES.items():
        kind = opts.get("kind", "scalar")
        fr = frame_of(w, h, seed, kw, opts)
        sec = ref.encode_dc_groups(fr, kind=kind)
        got = ref.decode_dc_groups(w, h, sec, kind=kind)
        # Lossless on what the format carries (one LUT id per transform).
        assert np.array_equal(got["dc"], fr.dc), name
        assert np.array_equal(got["ac_strategy"], fr.ac_strategy), name
        assert np.array_equal(got["quant_field"], fr.quant_field), name
        assert got["consumed"] == len(sec), name
        out[name + "/recipe"] = np.array(json.dumps(
            dict(w=w, h=h, seed=seed, kw=kw, opts=opts)))
        out[name + "/data"] = np.frombuffer(sec, np.uint8)
        for k in ("dc", "ac_strategy", "quant_field", "sigma_lut_ids"):
            out[name + "/" + k] = got[k]
        out[name + "/consumed"] = np.array(got["consumed"])
        print(name, len(sec), flush=True)
    np.savez_compressed(os.path.join(HERE, "dc_groups.npz"), **out)


if __name__ == "__main__":
    main()
