#!/usr/bin/env python3
"""Throughput of the decoder-side reconstruction on one MI355X.

Workload: one 4096x4096 (16.78 MP) synthetic frame at distance 1.0 with all
AC strategies (DCT8/4X4/2X2/16/32, identity, no-HF variants), predictions,
Gaborish k750, EPF + AR and dot patches. The frame (quantized coefficients
and side information, what the host holds after entropy decoding) is
uploaded once; each timed step is one full device-resident reconstruction
(DC dequant + DC EPF, LF predictions, dequant/CfL/HF prediction/IDCT, EPF +
AR, dots, Gaborish, OpsinToLinear) into the linear Image3F in HBM; the K
timed steps run back to back on the library's stream, the host
synchronised before the first and after the last.

Prints one JSON line. `roofline` covers the dominant kernel, timed with the
library's HIP events on the stream it launches on; `cpu_baseline` times the
reference's own AVX2 build (oracle/_ref, the reference's ThreadPool) on the
same frame on this host's cores.

Two further figures ride on the same line:
  batch    256 frames of seeded uneven sizes (W, H ~ U[256, 2048], distance
           ~ U[0.5, 3.0]; 32 distinct synthetic frames cycled) reconstructed
           by one graph launch per step (pikgpu_batch_run: one launch per
           phase over all frames, in the schedule pik_amd/autotune.py
           chose); MP/s.
  fwd_adj  forward model + adjoint (pikgpu_fwd_adj, device pointers) of a
           2048x2048 frame, loss sum (linear - target)^2, gradient with
           respect to the 256x256 float quant field and the dequantised
           coefficients; MP/s.
  quantize encoder-side quantization (pikgpu_quantize) of a 4096x4096 XYB
           image with the LF and HF prediction residuals (the reference
           encoder's default; a figure without them beside it): forward
           transforms, DC, residuals and dithered AC quantization; MP/s of
           the kernels (HIP events), and the rate with the host-to-device
           copy of the float image and the int16 downloads included.
  search   gradient quant-field search (pikgpu_search_quant_field) of 64
           1024x1024 images at Butteraugli distance 1.0, 20 Adam steps each
           from the reference encoder's initial field (committed fixture),
           one context and stream per image, wall time for all 64; beside
           it, on a sample of the same images, the reference's own search
           loop (its round-trip pieces, oracle/_ref) timed on the host's
           cores, and both results judged by the reference Butteraugli and
           sized by the reference's entropy coder.

The reconstruction figure also carries `with_copies_MPps` (the drop-in
call: frame upload from pinned host memory, run, linear download into
pinned memory) and, in the CPU-baseline leg, the deviation of the GPU output
from the reference's default AVX2 build (whose EPF divides with rcpps), the
host entropy decoding of the frame's AC groups and the parse of the whole
frame written as a bitstream by the reference encoder, each beside the
reference decoder on the same bytes, and the bitstream-to-image rate
(parse, upload, reconstruction).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("recon MP/s @16MP d=1.0 (1 GPU vs host cores); batched MP/s; "
          "fwd+adj MP/s")
HBM_PEAK_GBPS = 8000.0

# Algorithmic HBM bytes per pixel of the timed kernels (inputs read once,
# outputs written once; per-block side data amortised over 64 pixels).
#  epf_ar (KEpf on the image): idct in 3 x f32 + u8 guide in 3 x u8
#     + AR out 3 x f32 + quant field i32 + strategy u8 + sigma LUT id u8
#     per block
#  blocks_idct (KBlocks): int16 AC 3 x 2 B + idct out 3 x f32
#     + quant field i32, strategy u8, colour-tile data per block
#     + LLF, LF and pred2x2 f32 per block (1 + 4 + 4 values per channel)
#  final_opsin (KFinal): AR+dots image in 3 x f32 + linear out 3 x f32
BYTES_PER_PX = {
    "epf_ar": 12.0 + 3.0 + 12.0 + 6.0 / 64,
    "blocks_idct": 6.0 + 12.0 + 5.0 / 64 + 3 * 9 * 4.0 / 64,
    "final_opsin": 12.0 + 12.0,
}
KERNEL_NAMES = {"epf_ar": "KEpf", "blocks_idct": "KBlocks",
                "final_opsin": "KFinal"}
# Their rows in the traffic profile (the image EPF is KEpf<false>; the DC
# image's is KEpf<true>).
PROFILE_ROWS = {"epf_ar": "KEpf<false>", "blocks_idct": "KBlocks",
                "final_opsin": "KFinal"}
# Whole-path algorithmic bytes per pixel: int16 AC (6) + f32 DC per block
# (12/64) + i32 quant field per block (4/64) + strategy and sigma LUT id
# bytes per block (2/64) + f32 linear out (12).
PATH_BYTES_PER_PX = 6.0 + 12.0 / 64 + 4.0 / 64 + 2.0 / 64 + 12.0


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--dots", type=int, default=64)
    ap.add_argument("--adaptive-reconstruction", type=int, default=1,
                    choices=(0, 1),
                    help="frame header flag; 0 decodes without EPF/AR (the "
                    "library then moves the IDCT image with its plain copy "
                    "kernel, which calibrates the HBM byte counters)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="wall-time budget of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: host cores, "
                    "at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-frames", type=int, default=256,
                    help="frames per batch (0: skip the batch figure)")
    ap.add_argument("--batch-distinct", type=int, default=32,
                    help="distinct synthetic frames cycled over the batch")
    ap.add_argument("--adj-size", type=int, default=2048,
                    help="frame size of the fwd_adj figure (0: skip)")
    ap.add_argument("--quant-size", type=int, default=4096,
                    help="frame size of the quantize figure (0: skip)")
    ap.add_argument("--search-images", type=int, default=64,
                    help="1024x1024 images of the search figure (0: skip)")
    ap.add_argument("--search-distinct", type=int, default=16,
                    help="distinct search images cycled (<= 16)")
    ap.add_argument("--search-sample", type=int, default=2,
                    help="images the reference's search runs on")
    return ap.parse_args()


def host_threads():
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` on this workload, as recorded in
    profiles/hbm_traffic_recon_4096.md from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE) scaled by the copy-kernel calibration; None
    when no such record exists."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_recon_4096.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^\|\s*%s\s*\|(.*)\|\s*([0-9.eE+]+)\s*\|\s*$"
                     % re.escape(kernel))
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if m:
                return float(m.group(2))
    return None


VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4  # wave64 VALU instructions per second


def valu_roofline(kernel, avg_ms):
    """VALU issue rate of `kernel` against the chip's peak: instructions per
    wave x waves from profiles/sq_counters_recon_4096.md over the measured
    average time; None without the record."""
    path = os.path.join(ROOT, "profiles", "sq_counters_recon_4096.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^\|\s*%s\s*\|\s*4194304\s*\|\s*(\d+)\s*\|\s*(\d+)\s*\|"
                     r".*\|\s*([0-9.]+)%%\s*\|\s*[0-9.]+\s*\|\s*$" % kernel)
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if m:
                waves, per_wave = int(m.group(1)), int(m.group(2))
                insts = waves * per_wave
                rate = insts / (avg_ms * 1e-3)
                return {"instructions": insts,
                        "achieved_per_s": round(rate, 0),
                        "peak_per_s": VALU_ISSUE_PER_S,
                        "frac": round(rate / VALU_ISSUE_PER_S, 4),
                        "valu_busy": round(float(m.group(3)) / 100.0, 3)}
    return None


def cpu_baseline(frame, seconds, threads, gpu_out=None):
    """The reference's own reconstruction (AVX2 build, its ThreadPool) on
    the same frame, repeated until `seconds` have passed (at least 2
    frames). With the GPU's output, also how far the reference's default
    build (AVX2 EPF dividing with rcpps) lies from it: max relative error
    (|a - g| / max |g|) and the fraction of samples above 1e-4."""
    import numpy as np
    from oracle import ref
    if not ref.available("avx2"):
        return None
    avx2, _ = ref.recon(frame, threads=threads, stages=False, kind="avx2")
    dev = None
    if gpu_out is not None:
        d = np.abs(avx2.astype(np.float64) - gpu_out.astype(np.float64))
        rel = d / max(float(np.abs(gpu_out).max()), 1e-30)
        dev = {"max_rel": float(rel.max()),
               "frac_above_1e-4": float((rel > 1e-4).mean()),
               "bit_equal_frac": float((avx2.view(np.uint32) ==
                                        gpu_out.view(np.uint32)).mean())}
    reps, t0 = 0, time.perf_counter()
    while True:
        ref.recon(frame, threads=threads, stages=False, kind="avx2")
        reps += 1
        el = time.perf_counter() - t0
        if (reps >= 2 and el >= seconds) or el >= 2 * seconds:
            break
    mp = frame.xsize * frame.ysize / 1e6
    return {"value": round(mp * reps / el, 3), "unit": "MP/s",
            "cores": threads, "kind": "reference",
            "sample": "%d full %dx%d frames, reference AVX2 build "
                      "(DequantImageAC + ReconOpsinImage per group + "
                      "FinalizeFrameDecoding), %d threads, %.1f s"
                      % (reps, frame.xsize, frame.ysize, threads, el),
            "avx2_vs_gpu": dev}


def entropy_baseline(frame, threads, seconds=3.0):
    """Host entropy decoding of the frame's AC (pikgpu_decode_ac_pass)
    beside the reference decoder's own functions on the same bitstreams:
    the frame's 512x512 groups written by the reference encoder
    (EncodeToBitstream), decoded by the reference single-threaded
    (DecodeNoise, DecodeCoeffOrder, DecodeHistograms, DecodeAC per tile,
    GatherCoefficients; its pass loop runs groups on a thread pool) and by
    this library on 1 and on `threads` host threads. Output buffers are
    allocated once; each figure is the best of repeated runs."""
    import numpy as np
    from oracle import ref
    from pik_amd import bitstream
    if not ref.available("avx2"):
        return None
    groups = ref.encode_ac_groups(frame, kind="avx2")
    mp = frame.xsize * frame.ysize / 1e6
    by, bx = frame.ac_strategy.shape
    out = np.zeros((3, by, bx * 64), np.int16)
    want = ref.decode_ac_groups(frame, groups, check_dequant=False,
                                kind="avx2", out=out)[0].copy()

    def best(fn):
        t_end = time.perf_counter() + seconds
        b = 1e30
        while True:
            t = time.perf_counter()
            fn()
            b = min(b, time.perf_counter() - t)
            if time.perf_counter() > t_end:
                return b

    t_ref = best(lambda: ref.decode_ac_groups(frame, groups, False, "avx2",
                                              out))
    figures = {}
    for n in sorted({1, threads}):
        t = best(lambda: bitstream.decode_ac_pass(
            frame.xsize, frame.ysize, frame.ac_strategy, groups, threads=n,
            out=out))
        assert np.array_equal(out, want)
        figures["ours_%dt" % n] = round(mp / t, 1)
    return {"unit": "MP/s", "bytes": int(sum(len(g) for g in groups)),
            "groups": len(groups), "reference_1t": round(mp / t_ref, 1),
            **figures, "threads": threads,
            "over_reference_1t": round(figures["ours_1t"] * t_ref / mp, 2),
            "sample": "AC of the %dx%d frame, one pass, reference encoder "
                      "bitstreams (full histogram clustering); best of "
                      "%.0f s each; int16 output bit-equal to the "
                      "reference's" % (frame.xsize, frame.ysize, seconds)}


def frame_baseline(R, frame, threads, seconds=3.0):
    """The whole frame as a bitstream: written by the reference encoder's
    pieces (pass header, tables, quantizer, colour maps, DC groups, dots,
    dequant control field and map, group sizes, groups), parsed by the
    reference decoder's pieces single-threaded (the reference parses the
    pass-global sections on one thread and the groups on a pool) and by
    pikgpu_parse_frame on 1 and `threads` host threads, and decoded end to
    end on the GPU (pikgpu_decode_frame: parse, upload from pageable host
    memory, reconstruction; no download). The parsed frame is checked
    against the reference's parse. Best of repeated runs each."""
    import ctypes
    import numpy as np
    from oracle import ref
    from pik_amd import abi, bitstream
    if not ref.available("avx2"):
        return None
    data = ref.encode_frame(frame, kind="avx2")
    mp = frame.xsize * frame.ysize / 1e6
    bufs = {}
    want = ref.decode_frame(data, frame.xsize, frame.ysize, max_passes=1,
                            kind="avx2", reuse=bufs)
    got, info = bitstream.parse_frame(data, frame.xsize, frame.ysize,
                                      threads=threads)
    for k in ("dc", "ac_strategy", "quant_field", "sigma_lut_ids",
              "ytox_map", "ytob_map", "dequant_cf", "dequant_map"):
        assert np.array_equal(getattr(got, k), want[k]), k
    assert np.array_equal(got.ac, want["passes"][0])
    assert np.array_equal(got.dots, want["dots"])
    assert info["consumed"] == len(data) == want["consumed"]
    del got, info, want

    def best(fn):
        t_end = time.perf_counter() + seconds
        b = 1e30
        while True:
            t = time.perf_counter()
            fn()
            b = min(b, time.perf_counter() - t)
            if time.perf_counter() > t_end:
                return b

    t_ref = best(lambda: ref.decode_frame(data, frame.xsize, frame.ysize,
                                          max_passes=1, kind="avx2",
                                          reuse=bufs))
    del bufs
    lib = abi.load()
    raw = np.frombuffer(data, np.uint8)

    # the parsed frame is reused from call to call, as the reference's
    # buffers are
    h = ctypes.c_void_p()

    def parse(n):
        rc = lib.pikgpu_parse_frame(raw.ctypes.data, len(data), frame.xsize,
                                    frame.ysize, n, ctypes.byref(h), None, 0)
        assert rc == abi.OK

    figures = {}
    for n in sorted({1, threads}):
        figures["ours_%dt" % n] = round(mp / best(lambda: parse(n)), 1)
    lib.pikgpu_parsed_free(h)

    def decode():
        R.decode_frame(data, frame.xsize, frame.ysize, threads=threads)
        R.sync()

    t_e2e = best(decode)
    return {"unit": "MP/s", "bytes": len(data),
            "reference_1t": round(mp / t_ref, 1), **figures,
            "threads": threads,
            "over_reference_1t": round(figures["ours_1t"] * t_ref / mp, 2),
            "bitstream_to_image_MPps": round(mp / t_e2e, 1),
            "sample": "the %dx%d frame written whole by the reference "
                      "encoder's pieces (%d bytes); host parse of every "
                      "section into the frame the GPU takes, bit-equal to "
                      "the reference's; bitstream_to_image: parse on %d "
                      "threads + upload + reconstruction; best of %.0f s "
                      "each" % (frame.xsize, frame.ysize, len(data), threads,
                                seconds)}


def srgb8_f64(linear):
    """8-bit sRGB of a linear image in [0, 255] by the IEC 61966-2-1
    transfer function in float64, clamped, rounded half up (the reference's
    CastClip01 / ToExternal, pik/external_image.cc:190-197, 666-696)."""
    import numpy as np
    v = np.clip(linear.astype(np.float64) / 255.0, 0.0, 1.0)
    e = np.where(v <= 0.0031308, 12.92 * v,
                 1.055 * np.power(v, 1.0 / 2.4) - 0.055)
    return np.floor(np.clip(e, 0.0, 1.0) * 255.0 + 0.5).astype(np.int32)


def srgb8_check(R, linear):
    """The 8-bit output of the same run against the float64 transfer
    function of its linear output: max difference and mismatching-sample
    fraction."""
    import numpy as np
    R.set_output(3)
    R.run()
    got = R.download_srgb8().astype(np.int32).transpose(2, 0, 1)
    R.set_output(1)
    diff = np.abs(got - srgb8_f64(linear))
    return {"max_diff_lsb": int(diff.max()),
            "mismatch_frac": float((diff != 0).mean())}


def bench_with_copies(R, frame, steps):
    """MP/s of the drop-in call from host memory: upload (AC planes from
    pinned host memory), run, download of the linear image into pinned
    host memory, per frame."""
    import numpy as np
    from pik_amd.devmem import PinnedArray
    pac = PinnedArray.copy_of(frame.ac)
    pout = PinnedArray((3, frame.ysize, frame.xsize), np.float32)
    fr = frame.copy()
    fr.ac = pac.array
    R.upload(fr)
    R.run()
    R.download(out=pout.array)
    t0 = time.perf_counter()
    for _ in range(steps):
        R.upload(fr)
        R.run()
        R.download(out=pout.array)
    el = time.perf_counter() - t0
    return round(frame.xsize * frame.ysize / 1e6 * steps / el, 2)


def bench_per_image(frames, reps):
    """MP/s of the same frames decoded one after another on one context
    (upload once each, then run + sync per frame): the per-image launch
    sequence a batch replaces."""
    from pik_amd.runtime import Reconstructor
    mp = sum(f.xsize * f.ysize for f in frames) / 1e6
    with Reconstructor(0) as R:
        total = 0.0
        for f in frames:
            R.upload(f)
            R.run()
            R.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                R.run()
                R.sync()
            total += (time.perf_counter() - t0) / reps
    return round(mp / total, 2)


def _batch_frame(args):
    """Synthetic frame i of the batch (pik_amd.autotune's workload: size
    and distance drawn from the seeded generator)."""
    from pik_amd import autotune
    return autotune.workload_frame(*args)


def synth_batch(n, distinct, seed):
    """The batch's frames, synthesised by a process pool (run before the
    GPU is touched: the workers are forked)."""
    from multiprocessing import Pool
    t = time.perf_counter()
    distinct = max(1, min(distinct, n))
    with Pool(min(distinct, host_threads())) as pool:
        pool_frames = pool.map(_batch_frame,
                               [(i, seed) for i in range(distinct)])
    frames = [pool_frames[i % distinct] for i in range(n)]
    print("batch: %d frames (%d distinct), %.1f MP, synthesised in %.1f s"
          % (n, distinct, sum(f.xsize * f.ysize for f in frames) / 1e6,
             time.perf_counter() - t), file=sys.stderr, flush=True)
    return frames, distinct


def bench_batch(frames, distinct, steps, warmup):
    """MP/s of pikgpu_batch_run over the frames (device-resident, one graph
    launch per step)."""
    from pik_amd.runtime import BatchReconstructor
    n = len(frames)
    mp = sum(f.xsize * f.ysize for f in frames) / 1e6
    with BatchReconstructor(0) as B:
        tuned = B.load_schedule()
        sched = B.schedule()
        B.upload(frames)
        for _ in range(max(1, warmup)):
            B.run()
            B.sync()
        dev_ms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            B.run()
            B.sync()
            dev_ms += B.last_ms
        el = time.perf_counter() - t0
    return {"value": round(mp * steps / el, 2), "unit": "MP/s",
            "ms_per_step": round(el * 1000.0 / steps, 3),
            "device_ms_per_step": round(dev_ms / steps, 3),
            "frames": n, "distinct_frames": distinct,
            "megapixels": round(mp, 2), "steps": steps,
            "schedule": {"mode": ("queue", "branches")[sched["mode"]],
                         "order": ("upload", "largest first",
                                   "smallest first")[sched["order"]],
                         "persist": sched["persist"],
                         "source": "pik_amd/batch_schedule.json (autotuner)"
                                   if tuned else "library default"},
            "workload": "batch of %d frames, W,H ~ U[256,2048], d ~ "
                        "U[0.5,3.0], one graph launch per step" % n}


SEARCH_FIELDS = os.path.join(ROOT, "tests", "golden", "search_fields_1024.npz")


def _search_item(args):
    """Search image `seed`: the frame started from the reference encoder's
    initial field (the committed fixture) through SetQuantField's rule."""
    seed, field, qdc_f = args
    from pik_amd import synth
    fr, opsin, orig, lin = synth.make_search_inputs(1024, 1024, seed=seed,
                                                    num_rects=40, target=1.0)
    start = synth.max_over_transforms(field, fr.ac_strategy)
    fr.global_scale, fr.quant_dc, fr.quant_field = synth.set_quant_field(
        start, qdc_f)
    return fr, opsin, orig, lin, start


def search_inputs(distinct):
    """The first `distinct` search images (process pool, before the GPU is
    touched); None without the fixture."""
    import numpy as np
    from multiprocessing import Pool
    if distinct <= 0 or not os.path.exists(SEARCH_FIELDS):
        return None
    z = np.load(SEARCH_FIELDS)
    fields, qdc_f = z["fields"], float(z["quant_dc"])
    distinct = min(distinct, len(fields))
    t = time.perf_counter()
    with Pool(min(distinct, host_threads())) as pool:
        items = pool.map(_search_item,
                         [(i, fields[i], qdc_f) for i in range(distinct)])
    print("search: %d distinct 1024x1024 images synthesised in %.1f s"
          % (distinct, time.perf_counter() - t), file=sys.stderr, flush=True)
    return items


def bench_search(items, images, steps):
    """Wall time of the GPU search of `images` images (the distinct ones
    cycled), one context and stream each, 8 host threads."""
    from pik_amd.search import search_batch
    batch = [items[i % len(items)][:2] + (items[i % len(items)][3],)
             for i in range(images)]
    search_batch(batch[:2], threads=2, steps=2)  # warm
    t0 = time.perf_counter()
    res = search_batch(batch, threads=8, steps=steps)
    el = time.perf_counter() - t0
    mp = images * 1024 * 1024 / 1e6
    return {"value": round(mp / el, 2), "unit": "MP/s",
            "wall_s": round(el, 3), "ms_per_image": round(el * 1e3 / images, 2),
            "images": images, "distinct_images": len(items), "steps": steps,
            "workload": "quant-field search of %d 1024x1024 images at "
                        "Butteraugli distance 1.0, %d Adam steps each from "
                        "the reference's initial field, predictions off"
                        % (images, steps)}, res


def search_baseline(items, results, sample, threads):
    """The reference's own search loop (12 Butteraugli rounds, restated over
    its functions in oracle/_ref) on the first `sample` images, timed on the
    host's cores; both results judged by the reference Butteraugli and
    sized by the reference's entropy coder, beside the initial field's."""
    from oracle import ref
    from pik_amd.runtime import Reconstructor
    from pik_amd.search import apply_result
    if not ref.available("scalar"):
        return None
    rows = []
    with Reconstructor(0) as R:
        for k in range(min(sample, len(items))):
            fr, opsin, orig, lin, start = items[k]
            row = {}
            f0 = fr.copy()
            f0.ac, f0.dc = ref.quantize(fr, opsin)
            row["initial"] = {"distance": round(ref.butteraugli(
                orig, R.reconstruct(f0)), 4), "bytes": sum(
                ref.encoded_bytes(f0))}
            g = apply_result(fr, results[k])
            row["gpu"] = {"distance": round(ref.butteraugli(
                orig, R.reconstruct(g)), 4), "bytes": sum(
                ref.encoded_bytes(g))}
            fb = ref.find_best_quantization(fr, opsin, orig, start, 1.0,
                                            threads=threads)
            h = fr.copy()
            h.quant_field = fb["quant_field"]
            h.global_scale, h.quant_dc = fb["global_scale"], fb["quant_dc"]
            h.ac, h.dc = ref.quantize(h, opsin)
            row["reference"] = {"distance": round(fb["distance"], 4),
                                "bytes": sum(ref.encoded_bytes(h)),
                                "seconds": round(fb["seconds"], 3)}
            rows.append(row)
            print("search sample %d: %s" % (k, json.dumps(row)),
                  file=sys.stderr, flush=True)
    ref_s = sum(r["reference"]["seconds"] for r in rows) / len(rows)
    return {"images": rows, "reference_seconds_per_image": round(ref_s, 3),
            "threads": threads, "kind": "reference",
            "sample": "%d of the 1024x1024 images: the reference's search "
                      "loop (quantize + decode + Butteraugli x 12, oracle/"
                      "_ref), %d threads" % (len(rows), threads)}


# Algorithmic HBM bytes per pixel of fwd_adj: int16 AC in (6), target in
# (12), dL/dd out (12), plus per block the float and integer quant field,
# strategy byte and dL/dqf (13 B / 64).
ADJ_PATH_BYTES_PER_PX = 6.0 + 12.0 + 12.0 + 13.0 / 64


def bench_fwd_adj(size, steps, warmup, seed):
    """MP/s of pikgpu_fwd_adj on a size x size frame, device pointers."""
    import numpy as np
    from pik_amd import synth
    from pik_amd.devmem import DeviceArray
    from pik_amd.runtime import Reconstructor
    frame = synth.make_frame(size, size, seed=seed + 17, distance=1.0,
                             num_rects=100, num_dots=8)
    by, bx = frame.quant_field.shape
    with Reconstructor(0) as R:
        target = R.reconstruct(frame)
        R.upload(frame)
        rng = np.random.default_rng(seed)
        qf = DeviceArray.from_numpy((frame.quant_field * rng.uniform(
            0.8, 1.25, (by, bx))).astype(np.float32))
        tt = DeviceArray.from_numpy(target)
        dqf = DeviceArray((by, bx), np.float32)
        dc = DeviceArray((3, by, bx * 64), np.float32)
        loss = DeviceArray.zeros((1,), np.float64)
        for _ in range(max(1, warmup)):
            R.fwd_adj_device(qf, dqf, target=tt, dl_dcoeffs=dc, loss=loss)
            R.sync()
        fams = {}
        t0 = time.perf_counter()
        for _ in range(steps):
            R.fwd_adj_device(qf, dqf, target=tt, dl_dcoeffs=dc, loss=loss)
            R.sync()
            for k in ("adj_forward", "adj_backward"):
                fams[k] = fams.get(k, 0.0) + R.timings()[k]
        el = time.perf_counter() - t0
    mp = frame.xsize * frame.ysize / 1e6
    value = mp * steps / el
    return {"value": round(value, 2), "unit": "MP/s",
            "ms_per_step": round(el * 1000.0 / steps, 4),
            "kernel_ms": {k: round(v / steps, 4) for k, v in fams.items()},
            "path_GBps": round(ADJ_PATH_BYTES_PER_PX * value / 1e3, 2),
            "steps": steps,
            "workload": "fwd+adj %dx%d d=1.0, %dx%d float quant field, "
                        "loss sum (linear - target)^2, dL/dqf + dL/dd"
                        % (size, size, by, bx)}


def bench_quantize(size, steps, warmup, seed):
    """MP/s of pikgpu_quantize on a size x size XYB image, with the LF / HF
    prediction residuals (the reference encoder's default) and without."""
    from pik_amd import synth
    from pik_amd.runtime import Reconstructor
    frame, opsin = synth.make_opsin_frame(size, size, seed=seed + 29,
                                          distance=1.0, num_rects=100,
                                          predict=True)
    mp = frame.xsize * frame.ysize / 1e6
    out = {}
    with Reconstructor(0) as R:
        for predict in (1, 0):
            frame.predict_lf = frame.predict_hf = predict
            R.upload(frame)
            for _ in range(max(1, warmup)):
                R.quantize(opsin)
            kms = 0.0
            t0 = time.perf_counter()
            for _ in range(steps):
                R.quantize(opsin)
                kms += R.timings()["enc_quantize"]
            el = time.perf_counter() - t0
            if predict:
                out = {"value": round(mp * steps / (kms / 1e3), 2),
                       "unit": "MP/s", "kernel_ms": round(kms / steps, 4),
                       "with_copies_MPps": round(mp * steps / el, 2)}
            else:
                out["no_predictions"] = {
                    "value": round(mp * steps / (kms / 1e3), 2),
                    "kernel_ms": round(kms / steps, 4)}
    out.update({"steps": steps,
                "workload": "quantize %dx%d d=1.0 XYB image with LF and HF "
                            "prediction residuals (forward transforms, DC, "
                            "LF / HF residuals, dithered AC)" % (size, size)})
    return out


def main():
    a = parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the reconstruction of a frame "
              "runs on a single MI355X (multi-GPU sharding is not part of "
              "this build); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    batch = None
    if a.batch_frames > 0:
        batch = synth_batch(a.batch_frames, a.batch_distinct, a.seed)
    search_items = search_inputs(a.search_distinct) if a.search_images else None
    from pik_amd import synth
    from pik_amd.runtime import Reconstructor

    t = time.perf_counter()
    frame = synth.make_frame(a.size, a.size, seed=a.seed, distance=1.0,
                             num_rects=200, line_art=True, num_dots=a.dots)
    frame.adaptive_reconstruction = a.adaptive_reconstruction
    print("frame %dx%d synthesised in %.1f s" % (a.size, a.size,
          time.perf_counter() - t), file=sys.stderr, flush=True)
    mp = frame.xsize * frame.ysize / 1e6

    R = Reconstructor(0)
    R.upload(frame)
    for _ in range(a.warmup):
        R.run()
        R.sync()
    # Timed: K reconstructions back to back on the library's stream (each
    # runs whole, in stream order), the host synchronised before and after.
    R.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        R.run()
    R.sync()
    el = time.perf_counter() - t0
    ms_step = el * 1000.0 / a.steps
    # the last timed run's kernel-family times (HIP events on the library's
    # stream): the roofline's launch duration
    timed_ms = R.timings()
    # Per-family device times averaged: the same K runs again, each
    # synchronised so that its HIP events can be read (not timed).
    fams = {}
    for _ in range(a.steps):
        R.run()
        R.sync()
        for k, v in R.timings().items():
            fams[k] = fams.get(k, 0.0) + v
    fam_ms = {k: v / a.steps for k, v in fams.items()}
    value = mp * a.steps / el
    print("per-family ms/step: " + json.dumps(
        {k: round(v, 4) for k, v in fam_ms.items()}), file=sys.stderr,
        flush=True)

    # Dominant single-kernel family. `achieved` divides the path's
    # algorithmic bytes per pixel (int16 AC in, per-block side data, f32
    # linear out: PATH_BYTES_PER_PX) times the frame's pixels by that
    # kernel's average HIP-event time; the kernel's own input / output
    # bytes (BYTES_PER_PX) ride along as `kernel_bytes`. The kernel is
    # limited by VALU issue, not HBM: `valu` gives its issue rate from the
    # SQ counter record of the same workload (VALU instructions per wave x
    # waves) against the chip's peak (1024 SIMDs x 2.4 GHz, one wave64
    # instruction per 4 cycles).
    dom = max((k for k in BYTES_PER_PX if fam_ms.get(k, 0.0) > 0),
              key=lambda k: fam_ms[k])
    pad_px = (frame.xsize_blocks * 8) * (frame.ysize_blocks * 8)
    algo_bytes = PATH_BYTES_PER_PX * frame.xsize * frame.ysize
    kernel_bytes = BYTES_PER_PX[dom] * (pad_px if dom != "final_opsin"
                                        else frame.xsize * frame.ysize)
    launch_ms = timed_ms.get(dom) or fam_ms[dom]
    achieved = algo_bytes / (launch_ms * 1e-3) / 1e9
    roof = {"kernel": KERNEL_NAMES[dom], "bound": "hbm",
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic_from_profiles(PROFILE_ROWS[dom]),
            "algorithmic_bytes": int(algo_bytes),
            "kernel_bytes": int(kernel_bytes),
            "avg_ms": round(launch_ms, 4),
            "avg_ms_over_steps": round(fam_ms[dom], 4),
            "limiter": "valu",
            "valu": valu_roofline(KERNEL_NAMES[dom], launch_ms)}

    # the output, its 8-bit form, and the drop-in call with host copies
    gpu_out = R.download()
    srgb = srgb8_check(R, gpu_out)
    with_copies = bench_with_copies(R, frame, max(1, a.steps // 4))
    print("srgb8: %s; with copies %.1f MP/s" % (json.dumps(srgb),
          with_copies), file=sys.stderr, flush=True)

    cpu = None
    threads = a.cpu_threads or host_threads()
    entropy = frame_parse = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(frame, a.cpu_seconds, threads, gpu_out)
        entropy = entropy_baseline(frame, threads)
        print("host entropy: " + json.dumps(entropy), file=sys.stderr,
              flush=True)
        frame_parse = frame_baseline(R, frame, threads)
        print("host frame parse: " + json.dumps(frame_parse),
              file=sys.stderr, flush=True)
    acs = frame.ac_strategy
    first = (acs & 15) == 0
    names = {0: "DCT8", 1: "IDENTITY", 2: "DCT2X2", 3: "DCT4X4", 4: "DCT16",
             5: "DCT32", 6: "LINES", 7: "DCT8_NOHF", 8: "DCT4X4_NOHF"}
    import numpy as np
    hist = {names[r]: round(float(((acs >> 4) == r).mean()), 4)
            for r in range(9) if ((acs >> 4) == r).any()}

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "recon %dx%d d=1.0 single image, all AC "
                               "strategies, gaborish k750 + EPF/AR + dots"
                               % (a.size, a.size),
                   "megapixels": round(mp, 3), "seed": a.seed,
                   "adaptive_reconstruction": a.adaptive_reconstruction,
                   "path_GBps": round(PATH_BYTES_PER_PX * value / 1e3, 2),
                   "family_ms": {k: round(v, 4) for k, v in fam_ms.items()},
                   "with_copies_MPps": with_copies,
                   "srgb8": srgb,
                   "side_information": "generator's (AC strategy, quant "
                                       "field, colour correlation)",
                   "global_scale": int(frame.global_scale),
                   "quant_field_p5_p50_p95": [
                       int(v) for v in np.percentile(frame.quant_field,
                                                     [5, 50, 95])],
                   "block_fraction_by_strategy": hist,
                   "transforms": int(first.sum())},
        "roofline": roof,
        "cpu_baseline": cpu,
        "host_entropy": entropy,
        "host_frame_parse": frame_parse,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    R.close()
    if batch:
        out["batch"] = bench_batch(batch[0], batch[1], max(1, a.steps // 4),
                                   a.warmup)
        per = bench_per_image(batch[0][:batch[1]], 3)
        out["batch"]["per_image_MPps"] = per
        out["batch"]["over_per_image"] = round(out["batch"]["value"] / per, 3)
        out["batch"]["over_single_4096"] = round(out["batch"]["value"] /
                                                 value, 3)
        print("batch: " + json.dumps(out["batch"]), file=sys.stderr,
              flush=True)
    if search_items:
        out["search"], results = bench_search(search_items, a.search_images,
                                              20)
        print("search: " + json.dumps(out["search"]), file=sys.stderr,
              flush=True)
        if not a.no_cpu_baseline and a.search_sample > 0:
            sb = search_baseline(search_items, results, a.search_sample,
                                 threads)
            if sb:
                out["search"]["cpu_baseline"] = sb
                out["search"]["speedup_vs_reference"] = round(
                    sb["reference_seconds_per_image"] * a.search_images /
                    out["search"]["wall_s"], 2)
    if a.adj_size > 0:
        out["fwd_adj"] = bench_fwd_adj(a.adj_size, a.steps, a.warmup, a.seed)
        print("fwd_adj: " + json.dumps(out["fwd_adj"]), file=sys.stderr,
              flush=True)
    if a.quant_size > 0:
        out["quantize"] = bench_quantize(a.quant_size, max(1, a.steps // 4),
                                         a.warmup, a.seed)
        print("quantize: " + json.dumps(out["quantize"]), file=sys.stderr,
              flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
