"""Float64 statement of the 8-bit sRGB output: the IEC 61966-2-1 transfer
function of the [0, 1]-scaled linear sample, clamped to [0, 1], times 255,
rounded as int(v + 0.5) (pik/external_image.cc:190-197, 666-696)."""
import numpy as np


def srgb8(linear255):
    v = np.asarray(linear255, np.float64) / 255.0
    e = np.where(v <= 0.0031308, 12.92 * v,
                 1.055 * np.power(np.maximum(v, 0.0), 1.0 / 2.4) - 0.055)
    e = np.clip(e, 0.0, 1.0)
    return np.floor(e * 255.0 + 0.5).astype(np.int32)
