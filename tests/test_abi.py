"""C ABI checks that need no GPU: the library loads, exports every entry
point include/pikgpu.h declares, the ctypes mirror has the header's struct
layout, and host-side frame validation rejects undecodable frames with the
reference's notion of a decoding error."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import PikGpuError, validate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pikgpu.h")


def header_functions():
    text = open(HEADER).read()
    return re.findall(r"PIKGPU_API\s+[\w\s\*]+?\b(pikgpu_\w+)\s*\(", text)


def test_header_declares_expected_entry_points():
    names = header_functions()
    assert sorted(names) == sorted(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(abi, "_lib", None)
    monkeypatch.setattr(abi, "LIB_PATH", str(tmp_path / "libpikgpu.so"))
    with pytest.raises(RuntimeError, match="not built"):
        abi.load()


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "pik_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", text,
                                     re.M), fn
                assert "oracle/" not in text or fn == "build.py", fn


_LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "pikgpu.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("pikgpu_frame size %zu\n", sizeof(pikgpu_frame));
  printf("pikgpu_dot size %zu\n", sizeof(pikgpu_dot));
  printf("pikgpu_config size %zu\n", sizeof(pikgpu_config));
  F(pikgpu_frame, gaborish) F(pikgpu_frame, global_scale)
  F(pikgpu_frame, ac) F(pikgpu_frame, ac_stride) F(pikgpu_frame, dc)
  F(pikgpu_frame, quant_field) F(pikgpu_frame, ac_strategy)
  F(pikgpu_frame, sigma_lut_ids) F(pikgpu_frame, dequant_cf)
  F(pikgpu_frame, ytox_map) F(pikgpu_frame, ytob_map)
  F(pikgpu_frame, ytox_dc) F(pikgpu_frame, dequant_map)
  F(pikgpu_frame, dequant_table) F(pikgpu_frame, dequant_offsets)
  F(pikgpu_frame, num_dequant_offsets) F(pikgpu_frame, dots)
  F(pikgpu_frame, num_dots) F(pikgpu_frame, grayscale) F(pikgpu_frame, flags)
  F(pikgpu_dot, transform) F(pikgpu_dot, width) F(pikgpu_dot, pixels)
  printf("pikgpu_adj_io size %zu\n", sizeof(pikgpu_adj_io));
  F(pikgpu_adj_io, qf) F(pikgpu_adj_io, target) F(pikgpu_adj_io, in_stride)
  F(pikgpu_adj_io, dl_dqf) F(pikgpu_adj_io, dl_dcoeffs) F(pikgpu_adj_io, loss)
  printf("pikgpu_search_params size %zu\n", sizeof(pikgpu_search_params));
  F(pikgpu_search_params, learning_rate) F(pikgpu_search_params, qf_max)
  F(pikgpu_search_params, trust_high)
  printf("pikgpu_search_io size %zu\n", sizeof(pikgpu_search_io));
  F(pikgpu_search_io, target) F(pikgpu_search_io, weight)
  F(pikgpu_search_io, quant_field) F(pikgpu_search_io, ac)
  F(pikgpu_search_io, dc_stride) F(pikgpu_search_io, loss)
  F(pikgpu_search_io, lambda_used)
  printf("pikgpu_ac_pass size %zu\n", sizeof(pikgpu_ac_pass));
  F(pikgpu_ac_pass, ac_strategy) F(pikgpu_ac_pass, num_groups)
  F(pikgpu_ac_pass, group_size) F(pikgpu_ac_pass, ac)
  F(pikgpu_ac_pass, ac_stride) F(pikgpu_ac_pass, group_consumed)
  printf("pikgpu_dc_section size %zu\n", sizeof(pikgpu_dc_section));
  F(pikgpu_dc_section, data) F(pikgpu_dc_section, dc)
  F(pikgpu_dc_section, ac_strategy) F(pikgpu_dc_section, quant_field_stride)
  F(pikgpu_dc_section, sigma_lut_ids) F(pikgpu_dc_section, consumed)
  return 0;
}
"""


def test_ctypes_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE,
                         text=True).stdout
    types = {"pikgpu_frame": abi.FrameC, "pikgpu_dot": abi.Dot,
             "pikgpu_config": abi.ConfigC, "pikgpu_adj_io": abi.AdjIoC,
             "pikgpu_search_params": abi.SearchParamsC,
             "pikgpu_search_io": abi.SearchIoC,
             "pikgpu_ac_pass": abi.AcPassC,
             "pikgpu_dc_section": abi.DcSectionC}
    for line in out.splitlines():
        t, m, v = line.split()
        if m == "size":
            assert ctypes.sizeof(types[t]) == int(v), t
        else:
            assert getattr(types[t], m).offset == int(v), (t, m)


@pytest.fixture(scope="module")
def frame():
    return synth.make_frame(72, 56, seed=3, num_dots=2)


def test_valid_frame_passes(frame):
    validate(frame)


def _expect(frame, code, pattern, **changes):
    fr = frame.copy()
    for k, v in changes.items():
        setattr(fr, k, v)
    with pytest.raises(PikGpuError, match=pattern) as e:
        validate(fr)
    assert e.value.code == code


def test_rejects_bad_sizes_and_scalars(frame):
    _expect(frame, abi.EINVAL, "invalid quantizer", global_scale=0)
    _expect(frame, abi.EINVAL, "invalid quantizer", quant_dc=257)
    _expect(frame, abi.EINVAL, "gaborish", gaborish=5)
    _expect(frame, abi.EUNSUPPORTED, "fixed-sigma", epf_enable_adaptive=0)


def test_header_flags(frame):
    # FrameHeader::kNoise and kGradientMap change the decode; a frame with
    # either is refused rather than decoded without them
    _expect(frame, abi.EUNSUPPORTED, "kNoise", flags=abi.FLAG_NOISE)
    _expect(frame, abi.EUNSUPPORTED, "kGradientMap",
            flags=abi.FLAG_GRADIENT_MAP)
    _expect(frame, abi.EUNSUPPORTED, "kGradientMap",
            flags=abi.FLAG_GRADIENT_MAP | abi.FLAG_GRAYSCALE_OPT)
    _expect(frame, abi.EUNSUPPORTED, "unknown", flags=8)
    _expect(frame, abi.EUNSUPPORTED, "unknown", flags=1 << 31)
    fr = frame.copy()
    fr.flags = abi.FLAG_GRAYSCALE_OPT
    validate(fr)


def test_rejects_quant_field_and_lut(frame):
    qf = frame.quant_field.copy()
    qf[2, 3] = 0
    _expect(frame, abi.EINVAL, "quant field", quant_field=qf)
    qf[2, 3] = 257
    _expect(frame, abi.EINVAL, "quant field", quant_field=qf)
    lut = frame.sigma_lut_ids.copy()
    lut[0, 0] = 2
    _expect(frame, abi.EINVAL, "sigma lut", sigma_lut_ids=lut)


def test_rejects_inconsistent_strategies(frame):
    acs = np.zeros_like(frame.ac_strategy)
    acs[0, 0] = 0x9 << 4
    _expect(frame, abi.EINVAL, "invalid AC strategy", ac_strategy=acs)
    acs[0, 0] = abi.ACS_LINES << 4
    fr = frame.copy()
    fr.ac_strategy = acs.copy()
    validate(fr)  # LINES decodes (pik/ac_strategy.cc:680-724)
    acs[0, 0] = (abi.ACS_DCT << 4) | 1
    _expect(frame, abi.EINVAL, "block index", ac_strategy=acs)
    # a DCT16 missing its other three blocks
    acs[0, 0] = abi.ACS_DCT16X16 << 4
    _expect(frame, abi.EINVAL, "inconsistent", ac_strategy=acs)
    # a DCT16 starting on an odd block column
    acs[:] = 0
    acs[0:2, 1:3] = (abi.ACS_DCT16X16 << 4) | np.array([[0, 1], [2, 3]])
    _expect(frame, abi.EINVAL, "misaligned", ac_strategy=acs)
    # a DCT32 leaving the image (9 x 7 blocks)
    acs[:] = 0
    acs[4:7, 0:4] = (abi.ACS_DCT32X32 << 4) | np.arange(12).reshape(3, 4)
    _expect(frame, abi.EINVAL, "misaligned|inconsistent", ac_strategy=acs)


def _strategy_ok(acs):
    """Whole-map restatement of the strategy rule (AcStrategyImage::
    SetFromRaw, pik/ac_strategy.cc:1066-1099): every block names a valid
    raw strategy and index, and belongs to a complete, aligned transform
    inside the image whose blocks carry indices 0 .. s*s-1 in raster
    order."""
    by, bx = acs.shape
    for y in range(by):
        for x in range(bx):
            raw, idx = int(acs[y, x]) >> 4, int(acs[y, x]) & 15
            if raw > 8:
                return False
            s = synth.covered(raw)
            if idx >= s * s:
                return False
            oy, ox = y - idx // s, x - idx % s
            if oy < 0 or ox < 0 or oy % s or ox % s or oy + s > by or \
                    ox + s > bx:
                return False
            for k in range(s * s):
                if int(acs[oy + k // s, ox + k % s]) != (raw << 4) | k:
                    return False
    return True


def test_strategy_check_against_whole_map_rule(frame):
    """The library's check (each transform's blocks checked once, from its
    first block) accepts and rejects exactly what the whole-map rule does,
    on valid tilings with one or two entries changed."""
    rng = np.random.default_rng(7)
    by, bx = frame.ac_strategy.shape
    base = frame.ac_strategy
    kinds = [abi.ACS_DCT, abi.ACS_DCT16X16, abi.ACS_DCT32X32,
             abi.ACS_DCT4X4, abi.ACS_IDENTITY]
    seen = {True: 0, False: 0}
    for trial in range(300):
        acs = base.copy()
        for _ in range(int(rng.integers(1, 3))):
            y, x = int(rng.integers(0, by)), int(rng.integers(0, bx))
            raw = kinds[int(rng.integers(0, len(kinds)))]
            s = synth.covered(raw)
            acs[y, x] = (raw << 4) | int(rng.integers(0, s * s))
        want = _strategy_ok(acs)
        fr = frame.copy()
        fr.ac_strategy = acs
        try:
            validate(fr)
            got = True
        except PikGpuError as e:
            assert e.code == abi.EINVAL
            got = False
        assert got == want, trial
        seen[want] += 1
    assert seen[True] > 10 and seen[False] > 100, seen


def test_rejects_dequant_tables(frame):
    dm = frame.dequant_map.copy()
    dm[0, 0] = 200
    _expect(frame, abi.EINVAL, "missing table", dequant_map=dm)
    off = frame.dequant_offsets.copy()
    off[5] = frame.dequant_table.size
    _expect(frame, abi.EINVAL, "out of table", dequant_offsets=off)
    _expect(frame, abi.EINVAL, "multiple of 21",
            dequant_offsets=frame.dequant_offsets[:20].copy())
    cf = frame.dequant_cf.copy()
    cf[0, 0] = 16
    _expect(frame, abi.EINVAL, "control field", dequant_cf=cf)


def test_rejects_bad_dots(frame):
    dots = frame.dots.copy()
    dots[0]["xsize"] = 9
    _expect(frame, abi.EINVAL, "dot size", dots=dots)
    dots = frame.dots.copy()
    dots[0]["transform"] = 0
    dots[0]["x"] = 70
    _expect(frame, abi.EINVAL, "outside", dots=dots)
    dots = frame.dots.copy()
    dots[0]["transform"] = 1
    dots[0]["width"] = 0
    _expect(frame, abi.EINVAL, "degenerate", dots=dots)


def test_rejects_null_arrays_and_short_strides(frame):
    fc, keep = frame.to_c()
    lib = abi.load()
    msg = ctypes.create_string_buffer(128)
    fc.ac_stride = 8
    assert lib.pikgpu_validate_frame(ctypes.byref(fc), msg, 128) == abi.EINVAL
    assert b"stride" in msg.value
    fc, keep = frame.to_c()
    fc.quant_field = None
    assert lib.pikgpu_validate_frame(ctypes.byref(fc), msg, 128) == abi.EINVAL
    assert b"null" in msg.value
    assert lib.pikgpu_validate_frame(None, msg, 128) == abi.EINVAL
