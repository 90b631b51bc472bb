"""Pins of the float64 forward/adjoint model (oracle/adjoint.py) and of the
transform transposes the GPU adjoint uses (xform.h, host build).

  * The model's forward equals the oracle's float32 decode of the frame
    with HF prediction, adaptive reconstruction and dots off (the frame the
    model describes) to float rounding, for predict_lf on and off and
    Gaborish off / strengths 1 and 3: a dropped term, a wrong CfL sign or
    a transposed transform fails it.
  * The model's gradients equal central finite differences of its own
    forward (float64, relative step 1e-5): a wrong transpose fails it.
  * For every strategy, the host build of BlockAdjoint (the device code's
    forward transform with the norm scales, or IDENTITY's direct
    transpose) equals the explicit transpose of the float64 inverse
    transform matrix.
"""
import ctypes
import os

import numpy as np
import pytest

from pik_amd import abi, synth

from oracle import adjoint, port

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "pik_amd", "libpikgpu_host.so")

pytestmark = pytest.mark.skipif(not port.available(),
                                reason="oracle port not built")


def _frame(w, h, seed, predict_lf, gaborish):
    fr = synth.make_frame(w, h, seed=seed)
    fr.predict_lf = predict_lf
    fr.gaborish = gaborish
    return fr


CASES = [(72, 56, 5, 1, abi.GABORISH_750), (72, 56, 5, 0, abi.GABORISH_750),
         (130, 75, 7, 1, 0), (96, 64, 3, 1, 1), (40, 48, 9, 0, 3)]


@pytest.mark.parametrize("w,h,seed,plf,gab", CASES)
def test_model_forward_is_the_decode(w, h, seed, plf, gab):
    fr = _frame(w, h, seed, plf, gab)
    m = adjoint.Model(fr)
    lin, _ = m.forward(fr.quant_field.astype(np.float64))
    ref, _ = port.recon(adjoint.model_frame(fr), stages=False)
    assert lin.shape == ref.shape
    err = np.abs(lin - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


def test_model_gradient_matches_finite_differences():
    fr = _frame(72, 56, 5, 1, abi.GABORISH_750)
    m = adjoint.Model(fr)
    rng = np.random.default_rng(0)
    qf = fr.quant_field * rng.uniform(0.8, 1.25, fr.quant_field.shape)
    ref, _ = port.recon(adjoint.model_frame(fr), stages=False)
    target = ref + rng.normal(0, 0.5, ref.shape)
    _, _, dqf, dd = m.loss_and_grad(qf, target)

    def loss_qf(q):
        return ((m.forward(q)[0] - target) ** 2).sum()

    for (y, x) in [(0, 0), (3, 4), (6, 8), (2, 7), (6, 0)]:
        h = 1e-5 * qf[y, x]
        e = np.zeros_like(qf)
        e[y, x] = h
        fd = (loss_qf(qf + e) - loss_qf(qf - e)) / (2 * h)
        assert abs(fd - dqf[y, x]) <= 1e-5 * abs(dqf).max(), (y, x)

    d = m.dequant(qf)

    def loss_d(dv):
        return ((m.forward_d(dv)[0] - target) ** 2).sum()

    for (c, y, k) in [(1, 2, 70), (0, 3, 5), (2, 1, 200), (1, 4, 300),
                      (0, 6, 511)]:
        e = np.zeros_like(d)
        e[c, y, k] = 1e-4
        fd = (loss_d(d + e) - loss_d(d - e)) / 2e-4
        assert abs(fd - dd[c, y, k]) <= 1e-5 * abs(dd).max(), (c, y, k)
    # no effect at the LLF positions
    assert np.all(dd[:, m.llf_mask] == 0)


def _host():
    if not os.path.exists(HOST_LIB):
        pytest.skip("libpikgpu_host.so not built")
    L = ctypes.CDLL(HOST_LIB)
    L.pikgpu_host_block_adjoint.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p]
    L.pikgpu_host_adjoint_norms.argtypes = [ctypes.c_void_p]
    return L


@pytest.mark.parametrize("raw", [abi.ACS_DCT, abi.ACS_IDENTITY,
                                 abi.ACS_DCT2X2, abi.ACS_DCT4X4,
                                 abi.ACS_DCT16X16, abi.ACS_DCT32X32,
                                 abi.ACS_DCT_NOHF, abi.ACS_DCT4X4_NOHF])
def test_block_adjoint_is_the_transpose(raw):
    L = _host()
    n = 8 * adjoint.covered(raw)
    M = adjoint.transform_matrix(raw)
    rng = np.random.default_rng(raw)
    for _ in range(4):
        x = rng.normal(0, 1, n * n).astype(np.float32)
        out = np.zeros(n * n, np.float32)
        L.pikgpu_host_block_adjoint(raw, x.ctypes.data, out.ctypes.data)
        ref = M.T @ x.astype(np.float64)
        assert np.abs(out - ref).max() <= 3e-6 * np.abs(ref).max()


def test_adjoint_norms_closed_form():
    L = _host()
    d = np.zeros(60, np.float32)
    L.pikgpu_host_adjoint_norms(d.ctypes.data)
    off = {4: 0, 8: 4, 16: 12, 32: 28}
    for n, o in off.items():
        b = adjoint._basis(n)
        np.testing.assert_allclose(d[o:o + n], (b * b).sum(axis=1),
                                   rtol=1e-6)
