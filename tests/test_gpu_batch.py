"""Batched reconstruction (pikgpu_batch_*) in every schedule: one launch per
phase over all frames (queue; upload order, largest or smallest frame
first; one workgroup per work item or persistent workgroups looping over
them) and parallel per-frame branches. Each frame's output equals its
single-frame reconstruction bit for bit, and the single-frame path is held
to the oracle on a sample of the same frames."""
import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import (BatchReconstructor, Reconstructor,
                             batch_phase_can_persist, batch_phases)

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def batch():
    """32 frames, W, H in [256, 2048] (most small), distance in
    [0.5, 3.0], some with dots, some without AR, Gaborish or predictions,
    and their single-frame reconstructions."""
    rng = np.random.default_rng(41)
    frames = []
    for i in range(32):
        if i < 2:
            w, h = (2048, 300) if i == 0 else (264, 2048)
        else:
            w, h = (int(v) for v in rng.integers(256, 700, 2))
        fr = synth.make_frame(w, h, seed=200 + i, num_rects=8,
                              num_dots=int(i % 4 == 1) * 3,
                              distance=float(rng.uniform(0.5, 3.0)))
        if i % 7 == 3:
            fr.adaptive_reconstruction = 0
        if i % 9 == 4:
            fr.gaborish = 0
            fr.predict_lf = fr.predict_hf = 0
        if i % 11 == 5:
            fr.epf_use_sharpened = 1
        frames.append(fr)
    singles = []
    with Reconstructor(0) as r:
        for fr in frames:
            singles.append(r.reconstruct(fr))
    print("batch frames ready", flush=True)
    return frames, singles


def test_single_frames_match_oracle(batch):
    from oracle import port
    frames, singles = batch
    for i in (0, 3, 4, 5, 9, 17):
        print("oracle frame", i, frames[i].xsize, frames[i].ysize, flush=True)
        want, _ = port.recon(frames[i], stages=False)
        assert np.array_equal(_bits(singles[i]), _bits(want)), i


def _run(frames, **sched):
    with BatchReconstructor(0) as br:
        br.set_schedule(**sched)
        br.upload(frames)
        br.run()
        outs = [br.download(i) for i in range(len(frames))]
        br.run()  # the graph runs again
        br.sync()
        assert br.last_ms > 0
        again = br.download(len(frames) - 1)
    assert np.array_equal(_bits(again), _bits(outs[-1]))
    return outs


@pytest.mark.parametrize("order", [0, 1, 2])
def test_queue_orders(batch, order):
    frames, singles = batch
    print("queue order", order, flush=True)
    outs = _run(frames, mode=abi.BATCH_QUEUE, order=order)
    for i, (a, b) in enumerate(zip(outs, singles)):
        assert np.array_equal(_bits(a), _bits(b)), i


@pytest.mark.parametrize("k", [1, 4])
def test_queue_persistent(batch, k):
    frames, singles = batch
    print("queue persistent", k, flush=True)
    persist = [k if p else 0 for p in batch_phase_can_persist()]
    outs = _run(frames, mode=abi.BATCH_QUEUE, order=1, persist=persist)
    for i, (a, b) in enumerate(zip(outs, singles)):
        assert np.array_equal(_bits(a), _bits(b)), i


def test_branches(batch):
    frames, singles = batch
    print("branches", flush=True)
    outs = _run(frames[:12], mode=abi.BATCH_BRANCHES, order=0)
    for i, (a, b) in enumerate(zip(outs, singles[:12])):
        assert np.array_equal(_bits(a), _bits(b)), i


def test_bench_workload_frames():
    """The frames of the bench's batch figure (pik_amd.autotune's
    generator, seed 2: W, H ~ U[256, 2048], distance ~ U[0.5, 3.0]), 16
    distinct ones each twice, under the tuned schedule: every output equals
    the single-frame path's bit for bit, and the two largest equal the
    oracle's."""
    from oracle import port
    from pik_amd.autotune import workload_frame
    distinct = [workload_frame(i, 2) for i in range(16)]
    frames = distinct + distinct
    print("bench workload: %.1f MP" % (sum(f.xsize * f.ysize
                                              for f in frames) / 1e6),
          flush=True)
    with BatchReconstructor(0) as br:
        br.load_schedule()
        br.upload(frames)
        br.run()
        outs = [br.download(i) for i in range(len(frames))]
    with Reconstructor(0) as r:
        for i, fr in enumerate(distinct):
            want = r.reconstruct(fr)
            assert np.array_equal(_bits(outs[i]), _bits(want)), i
            assert np.array_equal(_bits(outs[i + 16]), _bits(want)), i
    big = sorted(range(16), key=lambda i: -distinct[i].xsize *
                 distinct[i].ysize)[:2]
    for i in big:
        print("oracle", distinct[i].xsize, distinct[i].ysize, flush=True)
        want, _ = port.recon(distinct[i], stages=False)
        assert np.array_equal(_bits(outs[i]), _bits(want)), i


@pytest.mark.parametrize("mode", [abi.BATCH_QUEUE, abi.BATCH_BRANCHES])
def test_multi_pass_frames(mode):
    """Frames with added AC passes (pikgpu_batch_add_pass, 1 to 3 passes
    per frame) decode as the single-frame path with the same passes does,
    bit for bit; a pass added after a run takes effect on the next run."""
    print("multi-pass batch, mode", mode, flush=True)
    frames = [synth.make_frame(w, h, seed=300 + k, num_rects=6)
              for k, (w, h) in enumerate([(264, 200), (520, 136), (72, 56),
                                          (300, 300)])]
    extra = [synth.make_extra_passes(fr, k % 3, seed=400 + k)
             for k, fr in enumerate(frames)]
    want = []
    with Reconstructor(0) as r:
        for fr, ex in zip(frames, extra):
            r.upload(fr)
            for p in ex:
                r.add_pass(p)
            r.run()
            want.append(r.download())
    with BatchReconstructor(0) as br:
        br.set_schedule(mode=mode, order=0)
        br.upload(frames)
        for k, ex in enumerate(extra):
            for p in ex:
                br.add_pass(k, p)
        br.run()
        for k in range(len(frames)):
            assert np.array_equal(_bits(br.download(k)), _bits(want[k])), k
        # one more pass to frame 0 after the run
        more = synth.make_extra_passes(frames[0], 1, seed=499)[0]
        br.add_pass(0, more)
        br.run()
        got0 = br.download(0)
        with pytest.raises(Exception):
            br.add_pass(len(frames), more)
    with Reconstructor(0) as r:
        r.upload(frames[0])
        for p in extra[0] + [more]:
            r.add_pass(p)
        r.run()
        assert np.array_equal(_bits(got0), _bits(r.download()))


def test_schedule_round_trip_and_refusals():
    with BatchReconstructor(0) as br:
        br.set_schedule(mode=abi.BATCH_QUEUE, order=2,
                        persist={"final": 3, "llf": 2})
        s = br.schedule()
        names = batch_phases()
        assert s["mode"] == abi.BATCH_QUEUE and s["order"] == 2
        assert s["persist"][names.index("final")] == 3
        assert s["persist"][names.index("llf")] == 2
        for bad in (dict(mode=5), dict(order=3), dict(persist=[17] * 19)):
            with pytest.raises(Exception):
                br.set_schedule(**bad)
        assert br.schedule() == s  # unchanged by the refusals
