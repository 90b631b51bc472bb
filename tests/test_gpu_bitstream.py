"""End to end from bitstreams: the committed AC group bitstreams and DC
groups sections (written by the reference encoder) decoded on the host by
pikgpu_decode_ac_pass / pikgpu_decode_dc_groups, the frame reconstructed on
the GPU, against the oracle's reconstruction of the frame the bitstreams
were written from, bit for bit."""
import json
import os

import numpy as np
import pytest

from pik_amd import bitstream, synth
from pik_amd.runtime import Reconstructor

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "ac_groups.npz")


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", ["tiny", "small_noise", "small_fast", "gray",
                                  "lines_groups", "groups_fast"])
def test_decode_then_reconstruct(rec, name):
    from oracle import port
    z = np.load(FIXTURE)
    r = json.loads(str(z[name + "/recipe"]))
    print("bitstream", name, r["w"], r["h"], flush=True)
    fr = synth.make_frame(r["w"], r["h"], seed=r["seed"], **r["kw"])
    assert np.array_equal(fr.ac_strategy, z[name + "/acs"])
    data = z[name + "/data"].tobytes()
    groups, pos = [], 0
    for s in z[name + "/sizes"]:
        groups.append(data[pos:pos + int(s)])
        pos += int(s)
    d = bitstream.decode_ac_pass(fr.xsize, fr.ysize, fr.ac_strategy, groups)
    want, _ = port.recon(fr, stages=False)
    decoded = fr.copy()
    decoded.ac = d["ac"]
    rec.upload(decoded)
    rec.run()
    assert np.array_equal(_bits(rec.download()), _bits(want))


DC_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                          "golden", "dc_groups.npz")


@pytest.mark.parametrize("name", ["tiny", "small_noise", "gray",
                                  "lines_groups", "groups_fast"])
def test_dc_and_ac_sections_then_reconstruct(rec, name):
    # Quantized DC, AC strategy, quant field and sigma LUT ids from the DC
    # groups section, AC from the AC group bitstreams, both decoded on the
    # host; the colour maps, dequant control field and header come from the
    # source frame (their sections are not decoded here).
    from oracle import port
    zd, za = np.load(DC_FIXTURE), np.load(FIXTURE)
    r = json.loads(str(zd[name + "/recipe"]))
    assert json.loads(str(za[name + "/recipe"]))["w"] == r["w"]
    print("dc+ac bitstreams", name, r["w"], r["h"], flush=True)
    fr = synth.make_frame(r["w"], r["h"], seed=r["seed"], **r["kw"])
    d = bitstream.decode_dc_groups(fr.xsize, fr.ysize,
                                   zd[name + "/data"].tobytes())
    data = za[name + "/data"].tobytes()
    groups, pos = [], 0
    for s in za[name + "/sizes"]:
        groups.append(data[pos:pos + int(s)])
        pos += int(s)
    a = bitstream.decode_ac_pass(fr.xsize, fr.ysize, d["ac_strategy"], groups)
    decoded = fr.copy()
    decoded.dc = d["dc"]
    decoded.ac_strategy = d["ac_strategy"]
    decoded.quant_field = d["quant_field"]
    decoded.sigma_lut_ids = d["sigma_lut_ids"]
    decoded.ac = a["ac"]
    want, _ = port.recon(fr, stages=False)
    rec.upload(decoded)
    rec.run()
    assert np.array_equal(_bits(rec.download()), _bits(want))


def test_dc_section_with_per_transform_luts(rec):
    # The bitstream carries one sigma LUT id per transform; the decoded
    # frame (first block's id everywhere in the transform) against the
    # oracle's decode of the same frame.
    from oracle import port
    zd = np.load(DC_FIXTURE)
    r = json.loads(str(zd["lines_lut/recipe"]))
    print("dc bitstream lines_lut", flush=True)
    fr = synth.make_frame(r["w"], r["h"], seed=r["seed"], **r["kw"])
    d = bitstream.decode_dc_groups(fr.xsize, fr.ysize,
                                   zd["lines_lut/data"].tobytes())
    decoded = fr.copy()
    decoded.sigma_lut_ids = d["sigma_lut_ids"]
    assert not np.array_equal(decoded.sigma_lut_ids, fr.sigma_lut_ids)
    want, _ = port.recon(decoded, stages=False)
    rec.upload(decoded)
    rec.run()
    assert np.array_equal(_bits(rec.download()), _bits(want))
