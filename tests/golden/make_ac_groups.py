"""Writes tests/golden/ac_groups.npz: AC group bitstreams written by the
reference encoder (EncodeToBitstream, pik/compressed_image.cc:457-573, once
per 512x512 group as EncodePass calls it) from pik_amd.synth frames, with
what the reference decoder makes of them: the int16 AC its own functions
decode in DecodeFromBitstream's order, the noise parameters and the bytes
read per group. The harness also runs DecodeFromBitstream<true> on every
group and checks that its dequantized coefficients equal DequantImageAC of
that int16, so the stored AC is what the reference decoder itself holds.

The cases cover a tile-sized frame, 16x16 / 32x32 / LINES mixes, several
groups with partial edge groups, noise parameters, the encoder's fast
(static context map) and full (clustered) histogram modes, a grayscale
frame, an all-zero AC frame and a frame of dense blocks with coefficients
up to +-32767. The reference code is the scalar build under oracle/_ref
(make -C oracle ref); run from the repository root:

    python3 tests/golden/make_ac_groups.py
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

# name: (w, h, seed, make_frame kwargs, encoder options)
CASES = {
    "tiny": (13, 7, 2, {}, {}),
    "small_noise": (200, 136, 4, {}, dict(noise=(0.012, 2.6, -0.004))),
    "small_fast": (200, 136, 4, {}, dict(fast_mode=True)),
    "gray": (264, 200, 11, dict(grayscale=True), {}),
    "lines_groups": (600, 520, 3, dict(lines=0.2), {}),
    "groups_fast": (1032, 600, 9, {}, dict(fast_mode=True)),
    "zero_ac": (136, 72, 5, {}, dict(craft="zero")),
    "dense_ac": (136, 72, 6, {}, dict(craft="dense")),
}


def craft(frame, kind, seed):
    """All-DCT8 frames with hand-made AC: all zero, or every other block
    dense with magnitudes up to 32767 (the widest the format codes)."""
    f = frame.copy()
    f.ac_strategy = np.zeros_like(f.ac_strategy)
    ac = np.zeros_like(f.ac)
    if kind == "dense":
        rng = np.random.default_rng(seed)
        by, bx = f.ac_strategy.shape
        blocks = ac.reshape(3, by, bx, 64)
        mag = rng.integers(-32767, 32768, size=blocks.shape)
        small = rng.integers(-3, 4, size=blocks.shape)
        pick = (np.arange(by)[:, None] + np.arange(bx)[None, :]) % 2 == 0
        blocks[:] = np.where(pick[None, :, :, None], mag, small)
        blocks[..., 0] = 0
    f.ac = ac
    return f


def uncoded_zeroed(frame):
    """The frame's AC with the slots no bitstream carries set to zero: slot 0
    of every block and, of 16x16 / 32x32 transforms, the S x S lowest
    frequencies (they come from the DC)."""
    ac = np.array(frame.ac, copy=True)
    by, bx = frame.ac_strategy.shape
    for y in range(by):
        for x in range(bx):
            a = int(frame.ac_strategy[y, x])
            s = {4: 2, 5: 4}.get(a >> 4, 1)
            if a & 15:
                continue
            for i in range(s):
                for j in range(s):
                    p = i * 8 * s + j  # coefficient (i, j) in the row-major
                    r, col = divmod(p, s * 64)  # S rows of S * 64
                    ac[:, y + r, x * 64 + col] = 0
    return ac


def main():
    out = {}
    for name, (w, h, seed, kw, enc) in CASES.items():
        fr = synth.make_frame(w, h, seed=seed, **kw)
        if "craft" in enc:
            fr = craft(fr, enc["craft"], seed)
        groups = ref.encode_ac_groups(fr, noise=enc.get("noise"),
                                      fast_mode=enc.get("fast_mode", False))
        ac, noise, used = ref.decode_ac_groups(fr, groups, check_dequant=True)
        # Entropy coding is lossless on the slots it carries.
        assert np.array_equal(ac, uncoded_zeroed(fr)), name
        out[name + "/recipe"] = np.array(json.dumps(
            dict(w=w, h=h, seed=seed, kw=kw, enc=enc)))
        out[name + "/acs"] = fr.ac_strategy
        out[name + "/data"] = np.frombuffer(b"".join(groups), np.uint8)
        out[name + "/sizes"] = np.array([len(g) for g in groups], np.int64)
        out[name + "/ac"] = ac
        out[name + "/noise"] = noise
        out[name + "/consumed"] = used
        print(name, [len(g) for g in groups], flush=True)
    np.savez_compressed(os.path.join(HERE, "ac_groups.npz"), **out)


if __name__ == "__main__":
    main()
