"""Whole frames from bitstream to image on the GPU: pikgpu_decode_frame
(host parse of every section, upload, every further pass, reconstruction)
on the committed frames written by the reference encoder
(tests/golden/frames.npz), against the oracle's reconstruction of the frame
the reference decoder read from the same bytes, bit for bit."""
import json
import os

import numpy as np
import pytest

from pik_amd import abi, tables
from pik_amd.frame import Frame
from pik_amd.runtime import PikGpuError, Reconstructor

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "frames.npz")
HEADER = ("flags", "gaborish", "predict_lf", "predict_hf",
          "adaptive_reconstruction", "epf_enable_adaptive", "epf_sigma",
          "epf_use_sharpened", "num_passes", "global_scale", "quant_dc",
          "ytox_dc", "ytob_dc", "num_table_sets", "num_dots", "frame_size")


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _reference_frame(z, name):
    """The frame as the reference decoder read it from the fixture's bytes
    (header fields, arrays, every pass), with the library tables once per
    encoded table set."""
    r = json.loads(str(z[name + "/recipe"]))
    h = dict(zip(HEADER, (int(v) for v in z[name + "/header"])))
    table, offsets = tables.default_tables()
    n = h["num_table_sets"]
    table_all = np.concatenate([table] * n)
    offs_all = np.concatenate([offsets + t * table.size for t in range(n)])
    passes = z[name + "/passes"]
    fr = Frame(
        xsize=r["w"], ysize=r["h"], ac=passes[0], dc=z[name + "/dc"],
        quant_field=z[name + "/quant_field"],
        ac_strategy=z[name + "/ac_strategy"],
        sigma_lut_ids=z[name + "/sigma_lut_ids"],
        dequant_cf=z[name + "/dequant_cf"], ytox_map=z[name + "/ytox_map"],
        ytob_map=z[name + "/ytob_map"], dequant_map=z[name + "/dequant_map"],
        dequant_table=table_all.astype(np.float32),
        dequant_offsets=offs_all.astype(np.uint32), dots=z[name + "/dots"],
        global_scale=h["global_scale"], quant_dc=h["quant_dc"],
        ytox_dc=h["ytox_dc"], ytob_dc=h["ytob_dc"], gaborish=h["gaborish"],
        predict_lf=h["predict_lf"], predict_hf=h["predict_hf"],
        adaptive_reconstruction=h["adaptive_reconstruction"],
        epf_enable_adaptive=h["epf_enable_adaptive"],
        epf_use_sharpened=h["epf_use_sharpened"],
        grayscale=int(bool(h["flags"] & abi.FLAG_GRAYSCALE_OPT)),
        flags=h["flags"]).contiguous()
    return fr, [passes[i] for i in range(1, len(passes))]


@pytest.mark.parametrize("name", ["tiny", "dots", "tdots", "passes", "gray",
                                  "tables", "plain", "sharp"])
def test_bitstream_to_image(rec, name):
    from oracle import port
    z = np.load(FIXTURE)
    data = z[name + "/data"].tobytes()
    fr, extra = _reference_frame(z, name)
    print("frame bitstream", name, fr.xsize, fr.ysize, len(extra) + 1,
          flush=True)
    want = port.recon_passes(fr, extra) if extra else \
        port.recon(fr, stages=False)[0]
    used = rec.decode_frame(data, fr.xsize, fr.ysize)
    assert used == len(data)
    got = rec.download()
    assert np.array_equal(_bits(got), _bits(want))


def test_refusals_leave_the_context_usable(rec):
    z = np.load(FIXTURE)
    data = z["tiny/data"].tobytes()
    print("frame refusals", flush=True)
    with pytest.raises(PikGpuError) as e:
        rec.decode_frame(data[:len(data) // 2], 13, 7)
    assert e.value.code == abi.EDATA
    rec.decode_frame(data, 13, 7)
    assert rec.download().shape == (3, 7, 13)
