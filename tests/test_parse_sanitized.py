"""The host frame parser under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/fuzz/fuzz_frame_parse.cc, compiled with
g++ together with the parser's sources (entropy, DC groups, frame parse,
library tables), parses every fixture frame of tests/golden/frames.npz
(bitstreams written by the reference encoder) and mutated copies of them
(bit flips, truncations, random bytes, header bytes); each must parse or be
rejected with PIKGPU_EDATA / PIKGPU_EUNSUPPORTED, with no sanitizer
report. Skipped where g++ or the sanitizer runtimes are missing."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pik_amd", "csrc")
SOURCES = ["entropy.cpp", "dc_groups.cpp", "frame_parse.cpp",
           "quant_tables.cpp", "host_tables.cpp"]


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fuzz")
    exe = str(d / "fuzz_frame_parse")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-msse4.1",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-Wno-unknown-pragmas",
           "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "fuzz", "fuzz_frame_parse.cc")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr or "ubsan" in r.stderr:
            pytest.skip("sanitizer runtimes not available")
        raise AssertionError(r.stderr[-2000:])
    z = np.load(os.path.join(ROOT, "tests", "golden", "frames.npz"))
    frames = []
    for name in sorted(set(k.split("/")[0] for k in z.keys())):
        recipe = json.loads(str(z[name + "/recipe"]))
        (d / (name + ".bin")).write_bytes(z[name + "/data"].tobytes())
        frames.append("%s:%d:%d" % (name, recipe["w"], recipe["h"]))
    return exe, str(d), frames


def test_fixture_frames_and_mutations_under_sanitizers(fuzzer):
    exe, d, frames = fuzzer
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:"
                            "verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, d, "300"] + frames, capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "fuzz done" in r.stdout
    parsed, rejected = [int(t) for t in r.stdout.split()[2:5:2]]
    assert parsed > 0 and rejected > 0
