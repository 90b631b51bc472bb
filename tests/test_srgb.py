"""The float64 sRGB statement used to check the GPU's 8-bit output, pinned
against the transfer function's published values (IEC 61966-2-1): black,
white, the linear-segment knee and the sRGB encodings of common levels."""
import numpy as np

from .srgb import srgb8


def test_endpoints_and_clamping():
    assert list(srgb8([0.0, 255.0, -10.0, 400.0])) == [0, 255, 0, 255]


def test_linear_segment():
    # below the knee (0.0031308) the curve is 12.92 x
    lin = 0.002 * 255
    assert srgb8([lin])[0] == int(12.92 * 0.002 * 255 + 0.5)


def test_known_levels():
    # 18% grey encodes to 0.4613 (118), 50% linear to 0.7354 (188); sRGB
    # 0.5 is linear 0.214041, so 0.2140 lands just below 127.5 (127)
    assert list(srgb8([0.18 * 255, 0.5 * 255, 0.2140 * 255])) == [118, 188, 127]


def test_round_trip_of_every_code():
    """Decoding each 8-bit code with the inverse curve and encoding it again
    gives the code back."""
    k = np.arange(256) / 255.0
    lin = np.where(k <= 0.04045, k / 12.92, ((k + 0.055) / 1.055) ** 2.4)
    assert np.array_equal(srgb8(lin * 255.0), np.arange(256))


def test_monotone():
    x = np.linspace(-5, 260, 10001)
    assert np.all(np.diff(srgb8(x)) >= 0)
