"""pikgpu_fwd_adj on the GPU against the float64 model (oracle/adjoint.py)
and against the decoder itself.

  * Forward: with qf equal to the integer quant field, `linear` is bit for
    bit pikgpu_run's output for the frame with HF prediction, adaptive
    reconstruction and dots off.
  * Gradient: dL/dqf and dL/dd against the float64 adjoint of the same
    model, tolerance 1e-4 of the gradient's norm (relative L2) and 1e-3 of
    its largest entry (max abs); the loss within 1e-5 relative.
  * Central finite differences of the GPU's own loss, relative step 1e-3,
    agree with its dL/dqf.
  * Target and dL/dlinear inputs, host and device pointers give the same
    bits; argument errors are rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

from pik_amd import abi, synth
from pik_amd.runtime import PikGpuError, Reconstructor

pytestmark = pytest.mark.gpu

GRAD_L2 = 1e-4
GRAD_MAX = 1e-3


@pytest.fixture(scope="module")
def rec():
    with Reconstructor(0) as r:
        yield r


def _model_frame(fr):
    f = fr.copy()
    f.predict_hf = 0
    f.adaptive_reconstruction = 0
    f.dots = f.dots[:0]
    return f


CASES = {
    "130x75_gab0": dict(w=130, h=75, seed=7, plf=1, gab=0),
    "96x64_nolf": dict(w=96, h=64, seed=3, plf=0, gab=abi.GABORISH_750),
    "264x200_groups": dict(w=264, h=200, seed=11, plf=1, gab=1),
    "13x7_sub_block": dict(w=13, h=7, seed=5, plf=1, gab=abi.GABORISH_750),
}


def _frame(c):
    fr = synth.make_frame(c["w"], c["h"], seed=c["seed"], num_dots=2)
    fr.predict_lf = c["plf"]
    fr.gaborish = c["gab"]
    return fr


@pytest.mark.parametrize("name", sorted(CASES))
def test_forward_is_bit_exact_decode(rec, name):
    print("case", name, flush=True)
    fr = _frame(CASES[name])
    want = rec.reconstruct(_model_frame(fr))
    rec.upload(fr)  # flags of the frame itself: the model ignores them
    zero = np.zeros_like(want)
    out = rec.fwd_adj(fr.quant_field.astype(np.float32), target=zero,
                      want_linear=True)
    assert np.array_equal(out["linear"].view(np.uint32),
                          want.view(np.uint32))
    ref_loss = float((want.astype(np.float64) ** 2).sum())
    assert abs(out["loss"] - ref_loss) <= 1e-5 * ref_loss
    # the decoder still runs afterwards
    rec.upload(_model_frame(fr))
    assert np.array_equal(rec.reconstruct(_model_frame(fr)), want)


def _rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()),
                                                  1e-30)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradient_matches_fp64_model(rec, name):
    from oracle import adjoint
    print("case", name, flush=True)
    fr = _frame(CASES[name])
    rng = np.random.default_rng(1)
    qf = (fr.quant_field * rng.uniform(0.8, 1.25, fr.quant_field.shape)
          ).astype(np.float32)
    base = rec.reconstruct(_model_frame(fr))
    target = (base + rng.normal(0, 0.5, base.shape)).astype(np.float32)
    rec.upload(fr)
    out = rec.fwd_adj(qf, target=target, want_linear=True, want_dcoeffs=True)
    m = adjoint.Model(fr)
    loss, lin, dqf, dd = m.loss_and_grad(qf.astype(np.float64),
                                         target.astype(np.float64))
    assert np.abs(out["linear"] - lin).max() <= 1e-4 * np.abs(lin).max()
    assert abs(out["loss"] - loss) <= 1e-5 * loss
    for got, ref in ((out["dl_dqf"], dqf), (out["dl_dcoeffs"], dd)):
        assert _rel_l2(got, ref) < GRAD_L2
        assert np.abs(got - ref).max() <= GRAD_MAX * np.abs(ref).max()
    assert np.all(out["dl_dcoeffs"][:, m.llf_mask] == 0)


def test_finite_differences_of_gpu_loss(rec):
    fr = _frame(CASES["96x64_nolf"])
    rng = np.random.default_rng(2)
    qf = (fr.quant_field * rng.uniform(0.8, 1.25, fr.quant_field.shape)
          ).astype(np.float32)
    base = rec.reconstruct(_model_frame(fr))
    target = (base + rng.normal(0, 2.0, base.shape)).astype(np.float32)
    rec.upload(fr)
    g = rec.fwd_adj(qf, target=target)["dl_dqf"]
    scale = np.abs(g).max()
    for (y, x) in [(0, 0), (3, 5), (7, 11), (4, 2), (7, 0)]:
        h = np.float32(1e-3) * qf[y, x]
        qp, qm = qf.copy(), qf.copy()
        qp[y, x] += h
        qm[y, x] -= h
        lp = rec.fwd_adj(qp, target=target)["loss"]
        lm = rec.fwd_adj(qm, target=target)["loss"]
        fd = (lp - lm) / float(qp[y, x] - qm[y, x])
        assert abs(fd - g[y, x]) <= 2e-2 * abs(g[y, x]) + 2e-3 * scale, \
            ((y, x), fd, g[y, x])


def test_upstream_gradient_and_device_pointers(rec):
    fr = _frame(CASES["264x200_groups"])
    rng = np.random.default_rng(3)
    qf = (fr.quant_field * rng.uniform(0.7, 1.4, fr.quant_field.shape)
          ).astype(np.float32)
    target = rng.normal(0, 20, (3, fr.ysize, fr.xsize)).astype(np.float32)
    rec.upload(fr)
    a = rec.fwd_adj(qf, target=target, want_linear=True, want_dcoeffs=True)
    up = ((a["linear"] - target) * np.float32(2)).astype(np.float32)
    b = rec.fwd_adj(qf, dl_dlinear=up, want_dcoeffs=True)
    assert np.array_equal(a["dl_dqf"], b["dl_dqf"])
    assert np.array_equal(a["dl_dcoeffs"], b["dl_dcoeffs"])
    # device pointers, asynchronous on the context stream
    from pik_amd.devmem import DeviceArray
    d_qf = DeviceArray.from_numpy(qf)
    d_target = DeviceArray.from_numpy(target)
    d_lin = DeviceArray((3, fr.ysize, fr.xsize), np.float32)
    d_dqf = DeviceArray(qf.shape, np.float32)
    d_dc = DeviceArray((3, fr.ysize_blocks, fr.xsize_blocks * 64), np.float32)
    d_loss = DeviceArray.zeros((1,), np.float64)
    rec.fwd_adj_device(d_qf, d_dqf, target=d_target, linear=d_lin,
                       dl_dcoeffs=d_dc, loss=d_loss)
    rec.sync()
    assert np.array_equal(d_lin.numpy(), a["linear"])
    assert np.array_equal(d_dqf.numpy(), a["dl_dqf"])
    assert np.array_equal(d_dc.numpy(), a["dl_dcoeffs"])
    assert d_loss.numpy()[0] == a["loss"]


def test_argument_errors(rec):
    fr = _frame(CASES["13x7_sub_block"])
    with Reconstructor(0) as fresh:
        fresh._frame = fr
        with pytest.raises(PikGpuError) as e:
            fresh.fwd_adj(fr.quant_field.astype(np.float32),
                          target=np.zeros((3, fr.ysize, fr.xsize), np.float32))
        assert e.value.code == abi.ESTATE
    rec.upload(fr)
    zero = np.zeros((3, fr.ysize, fr.xsize), np.float32)
    bad = fr.quant_field.astype(np.float32)
    bad[0, 0] = 0.0
    with pytest.raises(PikGpuError) as e:
        rec.fwd_adj(bad, target=zero)
    assert e.value.code == abi.EINVAL
    bad[0, 0] = np.nan
    with pytest.raises(PikGpuError):
        rec.fwd_adj(bad, target=zero)
    # raw calls: missing planes, partial output planes, loss without target
    lib = abi.load()
    qf = fr.quant_field.astype(np.float32)
    dqf = np.zeros_like(qf)
    io = abi.AdjIoC()
    io.qf = qf.ctypes.data
    io.qf_stride = fr.xsize_blocks
    io.dl_dqf = dqf.ctypes.data
    io.dqf_stride = fr.xsize_blocks
    io.in_stride = fr.xsize
    assert lib.pikgpu_fwd_adj(rec._ctx, ctypes.byref(io)) == abi.EINVAL
    io.dl_dlinear = (ctypes.c_void_p * 3)(*[zero[c].ctypes.data
                                            for c in range(3)])
    lin = np.zeros_like(zero)
    io.linear = (ctypes.c_void_p * 3)(lin[0].ctypes.data, None, None)
    io.linear_stride = fr.xsize
    assert lib.pikgpu_fwd_adj(rec._ctx, ctypes.byref(io)) == abi.EINVAL
    io.linear = (ctypes.c_void_p * 3)(None, None, None)
    loss = ctypes.c_double(0)
    io.loss = ctypes.addressof(loss)
    assert lib.pikgpu_fwd_adj(rec._ctx, ctypes.byref(io)) == abi.EINVAL
    io.loss = None
    io.in_stride = fr.xsize - 1
    assert lib.pikgpu_fwd_adj(rec._ctx, ctypes.byref(io)) == abi.EINVAL
    io.in_stride = fr.xsize
    assert lib.pikgpu_fwd_adj(rec._ctx, ctypes.byref(io)) == abi.OK


def test_lines_frames_are_refused(rec):
    # the model has no adjoint of the LINES transform: such a frame is
    # refused before any launch, and the context stays usable
    fr = _model_frame(synth.make_frame(64, 64, seed=9, lines=0.5))
    rec.upload(fr)
    qf = fr.quant_field.astype(np.float32)
    tgt = np.zeros((3, fr.ysize, fr.xsize), np.float32)
    with pytest.raises(PikGpuError) as e:
        rec.fwd_adj(qf, target=tgt)
    assert e.value.code == abi.EUNSUPPORTED
    ok = _model_frame(synth.make_frame(64, 64, seed=9))
    rec.upload(ok)
    rec.fwd_adj(ok.quant_field.astype(np.float32), target=tgt)


def test_full_size_2048_with_256_field(rec):
    """The forward + adjoint workload at its stated size: a 2048x2048 frame
    (256x256 quant field). The forward with the integer field is the
    decoder's output bit for bit at full size, and central finite
    differences of the GPU's own loss (relative step 1e-3) agree with
    dL/dqf on sampled blocks spread over the frame."""
    print("fwd+adj 2048x2048", flush=True)
    fr = synth.make_frame(2048, 2048, seed=19, distance=1.0, num_rects=100,
                          num_dots=8)
    assert fr.quant_field.shape == (256, 256)
    target = rec.reconstruct(fr)  # the full decode (HF prediction, AR, dots)
    want = rec.reconstruct(_model_frame(fr))
    rec.upload(fr)
    out = rec.fwd_adj(fr.quant_field.astype(np.float32), target=target,
                      want_linear=True)
    assert np.array_equal(out["linear"].view(np.uint32), want.view(np.uint32))
    assert out["loss"] > 0
    rng = np.random.default_rng(19)
    qf = (fr.quant_field * rng.uniform(0.8, 1.25, fr.quant_field.shape)
          ).astype(np.float32)
    g = rec.fwd_adj(qf, target=target)["dl_dqf"]
    scale = np.abs(g).max()
    assert np.isfinite(g).all() and scale > 0
    picks = [(0, 0), (255, 255), (0, 255), (128, 64)]
    picks += [tuple(int(v) for v in rng.integers(0, 256, 2)) for _ in range(8)]
    # and the blocks with the largest gradients
    big = np.argsort(np.abs(g).ravel())[-4:]
    picks += [divmod(int(i), 256) for i in big]
    for (y, x) in picks:
        h = np.float32(1e-3) * qf[y, x]
        qp, qm = qf.copy(), qf.copy()
        qp[y, x] += h
        qm[y, x] -= h
        lp = rec.fwd_adj(qp, target=target)["loss"]
        lm = rec.fwd_adj(qm, target=target)["loss"]
        fd = (lp - lm) / float(qp[y, x] - qm[y, x])
        assert abs(fd - g[y, x]) <= 2e-2 * abs(g[y, x]) + 2e-3 * scale, \
            ((y, x), fd, g[y, x])
