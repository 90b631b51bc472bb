"""Gradient quant-field search over many frames (pikgpu_search_quant_field).

The reference encoder searches the quant field of one image at a time with
twelve round trips judged by Butteraugli (FindBestQuantization,
pik/adaptive_quantization.cc:677-853). Here each frame's search runs on the
GPU (Adam on the log field, driven by the forward + adjoint of the decode);
a batch of frames gets one context and one stream per frame, driven from a
few host threads so that their kernels overlap on the device.

Per-block weights: the search equalizes w_b D_b, where D_b is block b's
squared linear-RGB error. Squared error in linear light overstates errors in
bright regions relative to what the eye sees (the response is close to a
cube root, as in the XYB transform, pik/opsin_image.cc:34-47); `block_weights`
scales each block by the square of that response's slope at the block's mean
luminance, (Y + 8)^(-4/3), normalized to mean 1.
"""
import threading

import numpy as np

from .runtime import Reconstructor

# Default search parameters (see pikgpu.h, pikgpu_search_params): the rate
# weight at 2.5 times the start's mean weighted distortion and each
# transform's field within [0.75, 1.25] of its start. On the 1024x1024
# benchmark images, started from the reference encoder's InitialQuantField,
# this ends within 0.2 % of the reference search's Butteraugli distance at
# 2-2.5 % fewer bytes; larger steps down (0.6 and below) let a few sensitive
# regions lose enough precision to raise the distance by 0.3-0.8.
DEFAULTS = dict(steps=20, learning_rate=0.05, lam=-2.5, trust=(0.75, 1.25))


def block_weights(linear):
    """Per-block weights (by, bx) float32 from the linear RGB original
    (3, H, W): ((Y_b + 8) / (mean Y + 8))^(-4/3) normalized to mean 1,
    Y_b the block's mean luminance."""
    lin = np.asarray(linear, np.float32)
    y = 0.2126 * lin[0] + 0.7152 * lin[1] + 0.0722 * lin[2]
    h, w = y.shape
    by, bx = -(-h // 8), -(-w // 8)
    yp = np.pad(y, ((0, by * 8 - h), (0, bx * 8 - w)), mode="edge")
    mean = yp.reshape(by, 8, bx, 8).mean(axis=(1, 3))
    wv = ((mean + 8.0) / (float(mean.mean()) + 8.0)) ** (-4.0 / 3)
    return (wv / wv.mean()).astype(np.float32)


def search_frame(rec, frame, opsin, linear, **params):
    """Uploads `frame` (its integer field is the start) and searches its
    quant field on the GPU; the context then holds the frame quantized with
    the result. Returns the dict of Reconstructor.search_quant_field."""
    p = dict(DEFAULTS)
    p.update(params)
    if "weight" not in p:
        p["weight"] = block_weights(linear)
    rec.upload(frame)
    return rec.search_quant_field(opsin, linear, **p)


def search_batch(items, device=0, threads=8, **params):
    """Searches every (frame, opsin, linear) of `items`, each on its own
    context and stream, from `threads` host threads (the library releases
    the interpreter while it waits for the device). Returns the results in
    order; each has the final integer field, AC and DC."""
    n = len(items)
    out = [None] * n
    errors = []
    lock = threading.Lock()
    next_i = [0]

    def worker():
        with Reconstructor(device) as rec:
            while True:
                with lock:
                    i = next_i[0]
                    next_i[0] += 1
                if i >= n or errors:
                    return
                fr, opsin, lin = items[i]
                try:
                    out[i] = search_frame(rec, fr, opsin, lin, **params)
                except Exception as e:  # reported after the join
                    with lock:
                        errors.append((i, e))
                    return

    pool = [threading.Thread(target=worker) for _ in range(max(1, threads))]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    if errors:
        i, e = errors[0]
        raise RuntimeError("frame %d: %s" % (i, e))
    return out


def apply_result(frame, result):
    """The frame with a search result's field and quantized coefficients."""
    f = frame.copy()
    f.quant_field = result["quant_field"].copy()
    f.ac = result["ac"].copy()
    f.dc = result["dc"].copy()
    return f
