"""End to end from bitstreams: the committed AC group bitstreams (written by
the reference encoder) decoded on the host by pikgpu_decode_ac_pass, the
frame reconstructed on the GPU, against the oracle's reconstruction of the
frame the bitstreams were written from, bit for bit."""
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
