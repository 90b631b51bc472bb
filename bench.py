#!/usr/bin/env python3
"""Throughput of the decoder-side reconstruction on one MI355X.

Workload: one 4096x4096 (16.78 MP) synthetic frame at distance 1.0 with all
AC strategies (DCT8/4X4/2X2/16/32, identity, no-HF variants), predictions,
Gaborish k750, EPF + AR and dot patches. The frame (quantized coefficients
and side information, what the host holds after entropy decoding) is
uploaded once; each timed step is one full device-resident reconstruction
(DC dequant + DC EPF, LF predictions, dequant/CfL/HF prediction/IDCT, EPF +
AR, dots, Gaborish, OpsinToLinear) into the linear Image3F in HBM.

Prints one JSON line. `roofline` covers the dominant kernel, timed with the
library's HIP events on the stream it launches on; `cpu_baseline` times the
reference's own AVX2 build (oracle/_ref, the reference's ThreadPool) on the
same frame on this host's cores.

Two further figures ride on the same line:
  batch    256 frames of seeded uneven sizes (W, H ~ U[256, 2048], distance
           ~ U[0.5, 3.0]; 32 distinct synthetic frames cycled) reconstructed
           by one graph launch per step (pikgpu_batch_run); MP/s.
  fwd_adj  forward model + adjoint (pikgpu_fwd_adj, device pointers) of a
           2048x2048 frame, loss sum (linear - target)^2, gradient with
           respect to the 256x256 float quant field and the dequantised
           coefficients; MP/s.
  quantize encoder-side quantization (pikgpu_quantize) of a 4096x4096 XYB
           image without predictions: forward transforms, DC and dithered
           AC quantization; MP/s of the kernels (HIP events), and the rate
           with the host-to-device copy of the float image and the int16
           downloads included.
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
    pat = re.compile(r"^\|\s*%s\s*\|(.*)\|\s*([0-9.eE+]+)\s*\|\s*$" % kernel)
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if m:
                return float(m.group(2))
    return None


def valu_busy_from_profiles(kernel):
    """VALU busy fraction of `kernel` on this workload, as recorded in
    profiles/sq_counters_recon_4096.md (rocprofv3 SQ counters); None when
    no such record exists."""
    path = os.path.join(ROOT, "profiles", "sq_counters_recon_4096.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^\|\s*%s\s*\|\s*4194304\s*\|.*\|\s*([0-9.]+)%%\s*\|"
                     r"\s*[0-9.]+\s*\|\s*$" % kernel)
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if m:
                return round(float(m.group(1)) / 100.0, 3)
    return None


def cpu_baseline(frame, seconds, threads):
    """The reference's own reconstruction (AVX2 build, its ThreadPool) on
    the same frame, repeated until `seconds` have passed (at least 2
    frames)."""
    from oracle import ref
    if not ref.available("avx2"):
        return None
    ref.recon(frame, threads=threads, stages=False, kind="avx2")  # warm
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
                      % (reps, frame.xsize, frame.ysize, threads, el)}


def _batch_frame(args):
    """Synthetic frame i of the batch: size and distance drawn from the
    seeded generator."""
    i, seed = args
    from pik_amd import synth
    import numpy as np
    rng = np.random.default_rng(seed * 1000003 + i)
    w, h = (int(v) for v in rng.integers(256, 2049, size=2))
    d = float(rng.uniform(0.5, 3.0))
    return synth.make_frame(w, h, seed=seed * 1000 + i, distance=d,
                            num_rects=40, num_dots=4)


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
            "workload": "batch of %d frames, W,H ~ U[256,2048], d ~ "
                        "U[0.5,3.0], one graph launch per step" % n}


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
    """MP/s of pikgpu_quantize on a size x size XYB image."""
    from pik_amd import synth
    from pik_amd.runtime import Reconstructor
    frame, opsin = synth.make_opsin_frame(size, size, seed=seed + 29,
                                          distance=1.0, num_rects=100,
                                          predict=False)
    with Reconstructor(0) as R:
        R.upload(frame)
        for _ in range(max(1, warmup)):
            R.quantize(opsin)
        kms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            R.quantize(opsin)
            kms += R.timings()["enc_quantize"]
        el = time.perf_counter() - t0
    mp = frame.xsize * frame.ysize / 1e6
    return {"value": round(mp * steps / (kms / 1e3), 2), "unit": "MP/s",
            "kernel_ms": round(kms / steps, 4),
            "with_copies_MPps": round(mp * steps / el, 2),
            "steps": steps,
            "workload": "quantize %dx%d d=1.0 XYB image, no predictions "
                        "(forward transforms + DC + dithered AC)"
                        % (size, size)}


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
    fams = {}
    R.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        R.run()
        R.sync()
        for k, v in R.timings().items():
            fams[k] = fams.get(k, 0.0) + v
    el = time.perf_counter() - t0
    ms_step = el * 1000.0 / a.steps
    fam_ms = {k: v / a.steps for k, v in fams.items()}
    value = mp * a.steps / el
    print("per-family ms/step: " + json.dumps(
        {k: round(v, 4) for k, v in fam_ms.items()}), file=sys.stderr,
        flush=True)

    # Dominant single-kernel family.
    dom = max((k for k in BYTES_PER_PX if fam_ms.get(k, 0.0) > 0),
              key=lambda k: fam_ms[k])
    pad_px = (frame.xsize_blocks * 8) * (frame.ysize_blocks * 8)
    algo_bytes = BYTES_PER_PX[dom] * (pad_px if dom != "final_opsin"
                                      else frame.xsize * frame.ysize)
    achieved = algo_bytes / (fam_ms[dom] * 1e-3) / 1e9
    roof = {"kernel": KERNEL_NAMES[dom], "bound": "hbm",
            "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic_from_profiles(KERNEL_NAMES[dom]),
            "algorithmic_bytes": int(algo_bytes),
            "avg_ms": round(fam_ms[dom], 4),
            # the EPF kernel is VALU-bound; its VALU busy fraction from the
            # SQ counters of the same workload
            "valu_busy": valu_busy_from_profiles(KERNEL_NAMES[dom])}

    cpu = None
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or host_threads()
        cpu = cpu_baseline(frame, a.cpu_seconds, threads)

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
                   "family_ms": {k: round(v, 4) for k, v in fam_ms.items()}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    R.close()
    if batch:
        out["batch"] = bench_batch(batch[0], batch[1], max(1, a.steps // 4),
                                   a.warmup)
        print("batch: " + json.dumps(out["batch"]), file=sys.stderr,
              flush=True)
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
