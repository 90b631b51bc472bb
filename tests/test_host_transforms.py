"""The block transforms and filter tables the HIP kernels use, checked on the
CPU through their host build (libpikgpu_host.so, same sources, same
-ffp-contract=off):

  - against the closed form of the transposed-scaled DCT the reference
    defines (pik/dct.h:351-507; pik/dct_test.cc:22-103 checks its fast
    transforms against the same slow form), in float64;
  - round trips TransformToPixels(TransformFromPixels(p)) == p for every
    strategy (pik/ac_strategy.cc:531-840);
  - bit for bit against the reference's own TransformToPixels /
    TransformFromPixels, EPF weights and Gaborish where oracle/_ref is built.
"""
import ctypes
import os

import numpy as np
import pytest

from pik_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "pik_amd", "libpikgpu_host.so")

STRATS = [abi.ACS_DCT, abi.ACS_IDENTITY, abi.ACS_DCT2X2, abi.ACS_DCT4X4,
          abi.ACS_DCT16X16, abi.ACS_DCT32X32, abi.ACS_DCT_NOHF,
          abi.ACS_DCT4X4_NOHF]
# strategies whose forward transform the inverse undoes exactly (LINES
# interpolates four pixels instead)
INVERTIBLE = STRATS
ALL = STRATS + [abi.ACS_LINES]


@pytest.fixture(scope="module")
def host():
    return ctypes.CDLL(HOST_LIB)


def size_of(raw):
    return {abi.ACS_DCT16X16: 16, abi.ACS_DCT32X32: 32}.get(raw, 8)


def host_tf(host, raw, direction, blk):
    n = size_of(raw)
    inp = np.ascontiguousarray(blk, np.float32).reshape(-1)
    out = np.zeros(n * n, np.float32)
    host.pikgpu_host_block_transform(raw, direction,
                                     inp.ctypes.data_as(ctypes.c_void_p),
                                     out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(n, n)


def slow_basis(n):
    """Row k: the k-th basis function of the transposed-scaled IDCT,
    w(k) cos(pi (2x+1) k / 2n) with w(0) = 1 and w(k) = 1/cos(k pi/16)
    (n = 8, the AAN scaling) or 2 cos(k pi/2n) (n = 4, 16, 32)."""
    k = np.arange(n)[:, None]
    x = np.arange(n)[None, :]
    if n == 8:
        w = np.where(k == 0, 1.0, 1.0 / np.cos(k * np.pi / 16))
    else:
        w = np.where(k == 0, 1.0, 2.0 * np.cos(k * np.pi / (2 * n)))
    return w * np.cos(np.pi * (2 * x + 1) * k / (2 * n))


def slow_idct2(coef):
    """pixels[y][x] = sum_{r,c} F[r][c] b_r(x) b_c(y): the coefficient row
    index is the horizontal frequency."""
    b = slow_basis(coef.shape[0])
    return b.T @ coef.T.astype(np.float64) @ b


@pytest.mark.parametrize("raw", [abi.ACS_DCT, abi.ACS_DCT16X16,
                                 abi.ACS_DCT32X32])
def test_dct_matches_closed_form(host, raw):
    n = size_of(raw)
    rng = np.random.default_rng(raw)
    for _ in range(20):
        coef = rng.uniform(-1, 1, (n, n)).astype(np.float32)
        got = host_tf(host, raw, 0, coef)
        want = slow_idct2(coef)
        assert np.abs(got - want).max() <= 2e-5 * max(1.0, np.abs(want).max())
        pix = rng.uniform(-1, 1, (n, n)).astype(np.float32)
        got = host_tf(host, raw, 1, pix)
        # forward = exact inverse of the closed-form IDCT
        want = np.linalg.solve(slow_basis(n).T,
                               np.linalg.solve(slow_basis(n).T,
                                               pix.astype(np.float64)).T).T
        want = want.T
        assert np.abs(slow_idct2(want) - pix).max() < 1e-9
        assert np.abs(got - want).max() <= 2e-6 * max(1.0, np.abs(want).max())


def test_dct8_dc_is_block_mean(host):
    rng = np.random.default_rng(5)
    pix = rng.uniform(0, 1, (8, 8)).astype(np.float32)
    coef = host_tf(host, abi.ACS_DCT, 1, pix)
    assert abs(coef[0, 0] - pix.mean()) < 1e-6


@pytest.mark.parametrize("raw", INVERTIBLE)
def test_round_trip(host, raw):
    n = size_of(raw)
    rng = np.random.default_rng(100 + raw)
    for _ in range(20):
        pix = rng.uniform(-2, 2, (n, n)).astype(np.float32)
        back = host_tf(host, raw, 0, host_tf(host, raw, 1, pix))
        assert np.abs(back - pix).max() < 2e-5


# LINES' four pixels on no diagonal line (row-major index) and the
# neighbours above and below them that they are interpolated from
# (pik/ac_strategy.cc:711-717).
LINES_GAPS = {21: (12, 30), 28: (19, 37), 35: (26, 44), 42: (33, 51)}


def test_lines_round_trip_and_gaps(host):
    rng = np.random.default_rng(7)
    on_lines = np.ones(64, bool)
    on_lines[list(LINES_GAPS)] = False
    for _ in range(20):
        pix = rng.uniform(-2, 2, (8, 8)).astype(np.float32)
        back = host_tf(host, abi.ACS_LINES, 0,
                       host_tf(host, abi.ACS_LINES, 1, pix)).reshape(64)
        assert np.abs(back[on_lines] - pix.reshape(64)[on_lines]).max() < 2e-5
        for g, (a, b) in LINES_GAPS.items():
            assert back[g] == np.float32((back[a] + back[b]) * np.float32(0.5))


def test_lines_constant_block(host):
    # a flat block: every line is constant, so only the slot-0 mean of the
    # four groups' Hadamard survives
    pix = np.full((8, 8), 3.5, np.float32)
    coef = host_tf(host, abi.ACS_LINES, 1, pix).reshape(64)
    assert coef[0] == np.float32(3.5)
    assert np.abs(np.delete(coef, 0)).max() < 1e-6


def test_dct4x4_is_four_dct4_with_hadamard_dc(host):
    """DCT4X4 (pik/ac_strategy.cc:590-613, 762-790): four 4-point 2-D DCTs
    interleaved with stride 2, the four DCs replaced by their 2x2 Hadamard
    / 4."""
    rng = np.random.default_rng(9)
    pix = rng.uniform(-1, 1, (8, 8)).astype(np.float32)
    coef = host_tf(host, abi.ACS_DCT4X4, 1, pix).astype(np.float64)
    b = slow_basis(4)
    sub = np.zeros((2, 2, 4, 4))
    for sy in range(2):
        for sx in range(2):
            p = pix[4 * sy:4 * sy + 4, 4 * sx:4 * sx + 4].astype(np.float64)
            f = np.linalg.solve(b.T, np.linalg.solve(b.T, p).T).T.T
            sub[sy, sx] = f
            for iy in range(4):
                for ix in range(4):
                    if iy or ix:
                        assert abs(coef[sy + 2 * iy, sx + 2 * ix] -
                                   f[iy, ix]) < 1e-5
    d = sub[:, :, 0, 0]
    had = np.array([d.sum(), d[0].sum() - d[1].sum(),
                    d[:, 0].sum() - d[:, 1].sum(),
                    d[0, 0] - d[0, 1] - d[1, 0] + d[1, 1]]) / 4
    assert np.allclose([coef[0, 0], coef[0, 1], coef[1, 0], coef[1, 1]],
                       had, atol=1e-6)


def test_identity_reproduces_pixels_from_residuals(host):
    """IDENTITY (pik/ac_strategy.cc:725-761): each 4x4 sub-block is its DC
    plus residuals against pixel (1, 1)."""
    rng = np.random.default_rng(3)
    pix = rng.uniform(-1, 1, (8, 8)).astype(np.float32)
    coef = host_tf(host, abi.ACS_IDENTITY, 1, pix)
    back = host_tf(host, abi.ACS_IDENTITY, 0, coef)
    assert np.abs(back - pix).max() < 1e-5
    # constant sub-blocks have zero residuals
    const = np.kron(np.array([[1.0, 2.0], [3.0, 4.0]]), np.ones((4, 4)))
    coef = host_tf(host, abi.ACS_IDENTITY, 1, const)
    mask = np.ones((8, 8), bool)
    mask[0:2, 0:2] = False
    assert np.abs(coef[mask]).max() == 0.0


# ---------------------------------------------------------------- reference

def _ref():
    from oracle import ref
    if not ref.available("scalar"):
        pytest.skip("reference harness not built (oracle/_ref)")
    return ref


@pytest.mark.parametrize("raw", ALL)
@pytest.mark.parametrize("direction", [0, 1])
def test_bit_exact_against_reference(host, raw, direction):
    ref = _ref()
    n = size_of(raw)
    rng = np.random.default_rng(1000 + 10 * raw + direction)
    for _ in range(50):
        blk = rng.uniform(-1, 1, (n, n)).astype(np.float32)
        want = ref.transform(raw, direction, blk)
        got = host_tf(host, raw, direction, blk)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _weight_from_table(mul, sad):
    """WeightFast's integer construction of the EPF weight
    (pik/epf_target.cc:304-376): the (negative) multiplier scales
    min(sad, 507) << 6, the product >> 16 is added to 127 << 7 (the upper
    16 bits of 1.0f) and the sum becomes the upper half of a float."""
    sad = np.minimum(sad, 507).astype(np.int64)
    v = ((sad << 6) * mul >> 16) + (127 << 7)
    return (v.astype(np.uint32) << 16).view(np.float32)


def test_mul_table_weights_match_reference(host):
    ref = _ref()
    table = np.zeros(1024, np.int32)
    host.pikgpu_host_mul_table(table.ctypes.data_as(ctypes.c_void_p))
    sads = np.arange(0, 600)
    for sigma in range(16, 673):
        want = np.array([ref.epf_weight(sigma, int(s)) for s in sads[::7]],
                        np.float32)
        got = _weight_from_table(int(table[sigma]), sads[::7])
        assert np.array_equal(got, want), sigma


def test_mul_table_weights_are_monotone(host):
    table = np.zeros(1024, np.int32)
    host.pikgpu_host_mul_table(table.ctypes.data_as(ctypes.c_void_p))
    sads = np.arange(0, 508)
    prev = None
    for sigma in range(16, 673):
        w = _weight_from_table(int(table[sigma]), sads)
        assert w[0] == 1.0 and np.all(np.diff(w) <= 0) and w.min() >= 0
        if prev is not None:  # larger sigma: flatter weights
            assert np.all(w >= prev)
        prev = w


# Unnormalised (w1, w2) of the four strengths, pik/gaborish.cc:16-85.
GABORISH_UNNORMALISED = {
    1: (0.056007960760453189, 0.045899074552453879),
    2: (0.75 * 0.13959942428275746, 0.75 * 0.074240717189152386),
    3: (0.875 * 0.11501538179658321, 0.875 * 0.089979079587015454),
    4: (0.11501538179658321, 0.089979079587015454),
}


def _gaborish_weights(host, strength):
    w = np.zeros(3, np.float32)
    host.pikgpu_host_gaborish_weights(strength,
                                      w.ctypes.data_as(ctypes.c_void_p))
    return w


@pytest.mark.parametrize("strength", [1, 2, 3, 4])
def test_gaborish_weights_normalised(host, strength):
    w = _gaborish_weights(host, strength)
    assert abs(float(w[0]) + 4 * float(w[1]) + 4 * float(w[2]) - 1) < 1e-6
    u1, u2 = GABORISH_UNNORMALISED[strength]
    assert abs(w[1] / w[0] - u1) < 1e-6 and abs(w[2] / w[0] - u2) < 1e-6


def _mirror_conv3(img, w):
    """3x3 symmetric convolution with mirrored borders (pik/convolve.h:
    502-640, Mirror: pik/image_ops.h:217-227), float64."""
    p = np.pad(img.astype(np.float64), ((0, 0), (1, 1), (1, 1)),
               mode="symmetric")
    c = p[:, 1:-1, 1:-1]
    edge = p[:, :-2, 1:-1] + p[:, 2:, 1:-1] + p[:, 1:-1, :-2] + p[:, 1:-1, 2:]
    corner = p[:, :-2, :-2] + p[:, :-2, 2:] + p[:, 2:, :-2] + p[:, 2:, 2:]
    return w[0] * c + w[1] * edge + w[2] * corner


@pytest.mark.parametrize("strength", [1, 2, 3, 4])
def test_gaborish_matches_reference_convolution(host, strength):
    ref = _ref()
    rng = np.random.default_rng(strength)
    img = rng.uniform(-1, 1, (3, 37, 29)).astype(np.float32)
    want = ref.gaborish(strength, img)
    got = _mirror_conv3(img, _gaborish_weights(host, strength).astype(
        np.float64))
    assert np.abs(got - want).max() < 1e-6
