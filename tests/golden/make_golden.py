"""Writes the golden reconstruction fixtures tests/golden/recon_*.npz and
the downsampled-decode fixtures tests/golden/downsample_*.npz.

Each fixture holds a synthetic frame (the decoder state after entropy
decoding, as pik_amd.frame.Frame fields) and what the reference's own
decode-side code computes from it: the DC after DequantDC and after
AdaptiveDCReconstruction, the ReconOpsinImage output (IDCT stage), the
AdaptiveReconstruction output and the final linear Image3F that
FinalizeFrameDecoding leaves. The reference code is the scalar build under
oracle/_ref (make -C oracle ref); run from the repository root:

    python3 tests/golden/make_golden.py [recon_<name> | downsample_<name> ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from pik_amd import abi, synth  # noqa: E402

# name -> (xsize, ysize, make_frame keyword arguments, header overrides)
CASES = {
    "small_default": (64, 64, dict(seed=1, num_dots=2), {}),
    "ragged_dots_lut": (136, 72, dict(seed=4, num_dots=5, lut_ids="rand"),
                        {}),
    "no_predict_g1000": (96, 88, dict(seed=7, predict=False),
                         dict(gaborish=abi.GABORISH_1000)),
    "no_ar_g500": (80, 48, dict(seed=9, num_dots=1),
                   dict(adaptive_reconstruction=0,
                        gaborish=abi.GABORISH_500)),
    "sharpened_d3": (72, 72, dict(seed=12, distance=3.0),
                     dict(epf_use_sharpened=1)),
    "grayscale": (104, 80, dict(seed=31, num_dots=2, grayscale=True), {}),
}
STAGES = ("dc_dequant", "dc", "idct", "ar")

# Downsampled decode (factors 2, 4, 8; frames without dots, as the
# reference requires): name -> (xsize, ysize, make_frame kwargs, overrides).
# Written to downsample_<name>.npz with the outputs ds2, ds4, ds8.
DS_CASES = {
    "default": (200, 136, dict(seed=21, num_dots=0), {}),
    "ragged_no_lf": (83, 45, dict(seed=22, num_dots=0),
                     dict(predict_lf=0)),
    "no_predict": (120, 72, dict(seed=23, predict=False, num_dots=0),
                   dict(gaborish=abi.GABORISH_1000)),
    "grayscale": (120, 88, dict(seed=32, num_dots=0, grayscale=True), {}),
}
DS_FACTORS = (2, 4, 8)


def make(name):
    xs, ys, kw, hdr = CASES[name]
    fr = synth.make_frame(xs, ys, **kw)
    for k, v in hdr.items():
        setattr(fr, k, v)
    out, st = ref.recon(fr, threads=1, stages=True, kind="scalar")
    return fr, out, st


def make_downsampled(name):
    xs, ys, kw, hdr = DS_CASES[name]
    fr = synth.make_frame(xs, ys, **kw)
    for k, v in hdr.items():
        setattr(fr, k, v)
    outs = {ds: ref.recon_downsampled(fr, ds, threads=1, kind="scalar")
            for ds in DS_FACTORS}
    return fr, outs


def main(only=()):
    """Writes every fixture, or those named in `only`."""
    for name in DS_CASES:
        if only and "downsample_" + name not in only:
            continue
        fr, outs = make_downsampled(name)
        path = os.path.join(HERE, "downsample_%s.npz" % name)
        d = {}
        for k, v in vars(fr).items():
            d["frame_" + k] = v if isinstance(v, np.ndarray) else np.array(v)
        for ds, out in outs.items():
            d["ds%d" % ds] = out
        np.savez_compressed(path, **d)
        print("wrote", path, os.path.getsize(path), "bytes")
    for name in CASES:
        if only and "recon_" + name not in only:
            continue
        fr, out, st = make(name)
        path = os.path.join(HERE, "recon_%s.npz" % name)
        d = {}
        for k, v in vars(fr).items():
            d["frame_" + k] = v if isinstance(v, np.ndarray) else np.array(v)
        for s in STAGES:
            d["stage_" + s] = st[s]
        d["linear"] = out
        np.savez_compressed(path, **d)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
