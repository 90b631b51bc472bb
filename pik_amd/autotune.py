"""Autotuner of the batched reconstruction's schedule (pikgpu_batch_schedule).

Measures, on the GPU it runs on, the device time of one batch of frames of
uneven sizes (W, H ~ U[256, 2048], distance ~ U[0.5, 3.0]: the batch
workload of bench.py) under each candidate schedule and writes the fastest
to pik_amd/batch_schedule.json, which BatchReconstructor.load_schedule()
applies:

  1. the mode and frame order: parallel per-frame branches, or one launch
     per phase over all frames (queue) in upload order, largest or smallest
     frame first, each with the phases that can loop run one workgroup per
     work item or k = 1, 2, 4 persistent workgroups per CU;
  2. per phase that can loop, k in {0, 1, 2, 4, 8} with the others fixed,
     kept when it shortens the batch by more than the run-to-run spread.

Run from the repository root on the GPU:

    python3 -m pik_amd.autotune [--frames 128] [--distinct 32] [--seed 2]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "batch_schedule.json")


def workload_frame(i, seed):
    """Frame i of the batch workload: size and distance from the seeded
    generator (the bench's batch figure uses the same frames)."""
    import numpy as np
    from . import synth
    rng = np.random.default_rng(seed * 1000003 + i)
    w, h = (int(v) for v in rng.integers(256, 2049, size=2))
    d = float(rng.uniform(0.5, 3.0))
    return synth.make_frame(w, h, seed=seed * 1000 + i, distance=d,
                            num_rects=40, num_dots=4)


def _frame_job(args):
    return workload_frame(*args)


def workload(n, distinct, seed, processes):
    """n frames cycling over `distinct` synthesised ones (a process pool;
    call before the GPU is touched: the workers are forked)."""
    from multiprocessing import Pool
    distinct = max(1, min(distinct, n))
    with Pool(max(1, min(distinct, processes))) as pool:
        made = pool.map(_frame_job, [(i, seed) for i in range(distinct)])
    return [made[i % distinct] for i in range(n)]


def measure(br, reps):
    """Median device ms of `reps` runs (after two warm-up runs)."""
    for _ in range(2):
        br.run()
        br.sync()
    times = []
    for _ in range(reps):
        br.run()
        br.sync()
        times.append(br.last_ms)
    times.sort()
    return times[len(times) // 2], times[-1] - times[0]


def tune(frames, reps=7, log=print):
    from . import abi
    from .runtime import (BatchReconstructor, batch_phase_can_persist,
                          batch_phases)
    names = batch_phases()
    loops = batch_phase_can_persist()
    mp = sum(f.xsize * f.ysize for f in frames) / 1e6
    table = []
    with BatchReconstructor(0) as br:
        br.upload(frames)

        def run(mode, order, persist):
            br.set_schedule(mode, order, persist)
            ms, spread = measure(br, reps)
            table.append({"mode": mode, "order": order,
                          "persist": list(persist), "ms": round(ms, 4),
                          "spread_ms": round(spread, 4),
                          "MPps": round(mp / (ms * 1e-3), 1)})
            log("  mode %d order %d persist %s: %.3f ms (%.0f MP/s)"
                % (mode, order, persist, ms, mp / (ms * 1e-3)))
            return ms, spread

        best = None
        cands = [(abi.BATCH_BRANCHES, 0, 0)]
        cands += [(abi.BATCH_QUEUE, o, k) for o in (0, 1, 2)
                  for k in (0, 1, 2, 4)]
        for mode, order, k in cands:
            persist = [k if p else 0 for p in loops]
            ms, spread = run(mode, order, persist)
            if best is None or ms < best[0]:
                best = (ms, spread, mode, order, persist)
        ms0, spread0, mode, order, persist = best
        if mode == abi.BATCH_QUEUE:
            for p, can in enumerate(loops):
                if not can:
                    continue
                for k in (0, 1, 2, 4, 8):
                    if k == persist[p]:
                        continue
                    trial = list(persist)
                    trial[p] = k
                    ms, spread = run(mode, order, trial)
                    if ms < ms0 - max(spread0, spread, 0.01 * ms0):
                        ms0, spread0, persist = ms, spread, trial
    return {"mode": mode, "order": order, "persist": persist,
            "phases": names, "ms": round(ms0, 4),
            "MPps": round(mp / (ms0 * 1e-3), 1),
            "frames": len(frames), "megapixels": round(mp, 2),
            "table": table}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--frames", type=int, default=128)
    ap.add_argument("--distinct", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--reps", type=int, default=7)
    ap.add_argument("--out", default=OUT)
    a = ap.parse_args()
    t = time.perf_counter()
    frames = workload(a.frames, a.distinct, a.seed, os.cpu_count() or 1)
    print("workload: %d frames (%d distinct), %.1f MP, %.1f s"
          % (len(frames), a.distinct,
             sum(f.xsize * f.ysize for f in frames) / 1e6,
             time.perf_counter() - t), flush=True)
    res = tune(frames, a.reps,
               log=lambda s: print(s, flush=True))
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print("chosen: mode %d order %d persist %s: %.3f ms, %.0f MP/s -> %s"
          % (res["mode"], res["order"], res["persist"], res["ms"],
             res["MPps"], a.out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
