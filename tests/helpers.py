"""Shared helpers of the test suite: golden fixtures and the checker."""
import glob
import os

import numpy as np

from pik_amd.frame import Frame

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golden_names():
    return sorted(os.path.basename(p)[len("recon_"):-len(".npz")]
                  for p in glob.glob(os.path.join(GOLDEN, "recon_*.npz")))


def load_golden(name):
    """Returns (Frame, {"dc_dequant", "dc", "idct", "ar", "linear"})."""
    z = np.load(os.path.join(GOLDEN, "recon_%s.npz" % name),
                allow_pickle=False)
    kw = {}
    for k in z.files:
        if k.startswith("frame_"):
            v = z[k]
            kw[k[len("frame_"):]] = v if v.ndim else v.item()
    want = {k[len("stage_"):]: z[k] for k in z.files if k.startswith("stage_")}
    want["linear"] = z["linear"]
    return Frame(**kw).contiguous(), want


def downsample_names():
    return sorted(os.path.basename(p)[len("downsample_"):-len(".npz")]
                  for p in glob.glob(os.path.join(GOLDEN, "downsample_*.npz")))


def load_downsample(name):
    """Returns (Frame, {factor: (3, ceil(H / factor), ceil(W / factor))})
    of a downsampled-decode fixture."""
    z = np.load(os.path.join(GOLDEN, "downsample_%s.npz" % name),
                allow_pickle=False)
    kw = {}
    for k in z.files:
        if k.startswith("frame_"):
            v = z[k]
            kw[k[len("frame_"):]] = v if v.ndim else v.item()
    want = {int(k[2:]): z[k] for k in z.files if k.startswith("ds")}
    return Frame(**kw).contiguous(), want


def checker_recon(frame, stages=True, threads=1):
    """The oracle's reconstruction of `frame`: the reference's own code
    (scalar build under oracle/_ref) where it is built, else the C
    restatement. Returns (linear, stages dict, kind)."""
    from oracle import ref
    if ref.available("scalar"):
        out, st = ref.recon(frame, threads=threads, stages=stages)
        return out, st, "reference"
    from oracle import port
    if port.available():
        out, st = port.recon(frame, stages=stages)
        return out, st, "port"
    raise RuntimeError("no oracle built: run __graft_entry__.build()")


def max_rel(got, want):
    """Max abs difference relative to the max magnitude of `want`."""
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    return float(d.max() / max(float(np.abs(want).max()), 1e-30))


def bit_equal_fraction(got, want):
    g = np.ascontiguousarray(got, np.float32).view(np.uint32)
    w = np.ascontiguousarray(want, np.float32).view(np.uint32)
    return float((g == w).mean())
