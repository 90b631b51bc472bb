// Per-strategy block transforms shared by the device kernels and the host
// self-check library: the direct 8x8 forms (DCT2X2, IDENTITY) and whole-block
// transposed-scaled transforms built on dct1d.h.
#pragma once

#include "common.h"
#include "dct1d.h"

namespace pikgpu {

// Direct 8x8 inverse transforms: DCT2X2 pyramid (pik/ac_strategy.cc:
// 791-803, 330-356), IDENTITY (pik/ac_strategy.cc:725-761), and LINES is not
// emitted by the encoder (pik/ac_strategy.cc:18) and decodes to zeros here.
PIKGPU_HD void DirectInverse8(int raw, const float* c, float* p,
                                      int pstride) {
  if (raw == kDCT2X2) {
    float a[64];
    for (int i = 0; i < 64; ++i) a[i] = c[i];
    for (int S = 2; S <= 8; S *= 2) {
      const int h = S / 2;
      float t[64];
      for (int y = 0; y < h; ++y) {
        for (int x = 0; x < h; ++x) {
          const float c00 = a[y * 8 + x];
          const float c01 = a[y * 8 + h + x];
          const float c10 = a[(y + h) * 8 + x];
          const float c11 = a[(y + h) * 8 + h + x];
          t[(2 * y) * 8 + 2 * x] = c00 + c01 + c10 + c11;
          t[(2 * y) * 8 + 2 * x + 1] = c00 + c01 - c10 - c11;
          t[(2 * y + 1) * 8 + 2 * x] = c00 - c01 + c10 - c11;
          t[(2 * y + 1) * 8 + 2 * x + 1] = c00 - c01 - c10 + c11;
        }
      }
      for (int y = 0; y < S; ++y) {
        for (int x = 0; x < S; ++x) a[y * 8 + x] = t[y * 8 + x];
      }
    }
    for (int y = 0; y < 8; ++y) {
      for (int x = 0; x < 8; ++x) p[y * pstride + x] = a[y * 8 + x];
    }
    return;
  }
  if (raw == kIdentity) {
    const float b00 = c[0], b01 = c[1], b10 = c[8], b11 = c[9];
    float dcs[4];
    dcs[0] = b00 + b01 + b10 + b11;
    dcs[1] = b00 + b01 - b10 - b11;
    dcs[2] = b00 - b01 + b10 - b11;
    dcs[3] = b00 - b01 - b10 + b11;
    for (int y = 0; y < 2; ++y) {
      for (int x = 0; x < 2; ++x) {
        const float block_dc = dcs[y * 2 + x];
        float residual_sum = 0;
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            if (ix == 0 && iy == 0) continue;
            residual_sum += c[(y + iy * 2) * 8 + x + ix * 2];
          }
        }
        const float center = block_dc - residual_sum * (1.0f / 16);
        p[(4 * y + 1) * pstride + 4 * x + 1] = center;
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            if (ix == 1 && iy == 1) continue;
            p[(y * 4 + iy) * pstride + x * 4 + ix] =
                c[(y + iy * 2) * 8 + x + ix * 2] + center;
          }
        }
        p[y * 4 * pstride + x * 4] = c[(y + 2) * 8 + x + 2] + center;
      }
    }
    return;
  }
  for (int y = 0; y < 8; ++y) {
    for (int x = 0; x < 8; ++x) p[y * pstride + x] = 0.0f;
  }
}

// Direct 8x8 forward transforms (pik/ac_strategy.cc:569-589, 614-618).
PIKGPU_HD void DirectForward8(int raw, const float* p, int pstride,
                                      float* c) {
  if (raw == kDCT2X2) {
    float a[64];
    for (int y = 0; y < 8; ++y) {
      for (int x = 0; x < 8; ++x) a[y * 8 + x] = p[y * pstride + x];
    }
    for (int S = 8; S >= 2; S /= 2) {
      const int h = S / 2;
      float t[64];
      for (int y = 0; y < h; ++y) {
        for (int x = 0; x < h; ++x) {
          const float c00 = a[(2 * y) * 8 + 2 * x];
          const float c01 = a[(2 * y) * 8 + 2 * x + 1];
          const float c10 = a[(2 * y + 1) * 8 + 2 * x];
          const float c11 = a[(2 * y + 1) * 8 + 2 * x + 1];
          float r00 = c00 + c01 + c10 + c11;
          float r01 = c00 + c01 - c10 - c11;
          float r10 = c00 - c01 + c10 - c11;
          float r11 = c00 - c01 - c10 + c11;
          r00 *= 0.25f;
          r01 *= 0.25f;
          r10 *= 0.25f;
          r11 *= 0.25f;
          t[y * 8 + x] = r00;
          t[y * 8 + h + x] = r01;
          t[(y + h) * 8 + x] = r10;
          t[(y + h) * 8 + h + x] = r11;
        }
      }
      for (int y = 0; y < S; ++y) {
        for (int x = 0; x < S; ++x) a[y * 8 + x] = t[y * 8 + x];
      }
    }
    for (int i = 0; i < 64; ++i) c[i] = a[i];
    return;
  }
  if (raw == kIdentity) {
    for (int i = 0; i < 64; ++i) c[i] = 0.0f;
    for (int y = 0; y < 2; ++y) {
      for (int x = 0; x < 2; ++x) {
        float block_dc = 0;
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            block_dc += p[(y * 4 + iy) * pstride + x * 4 + ix];
          }
        }
        block_dc *= 1.0f / 16;
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            if (ix == 1 && iy == 1) continue;
            c[(y + iy * 2) * 8 + x + ix * 2] =
                p[(y * 4 + iy) * pstride + x * 4 + ix] -
                p[(y * 4 + 1) * pstride + x * 4 + 1];
          }
        }
        c[(y + 2) * 8 + x + 2] = c[y * 8 + x];
        c[y * 8 + x] = block_dc;
      }
    }
    const float b00 = c[0], b01 = c[1], b10 = c[8], b11 = c[9];
    c[0] = (b00 + b01 + b10 + b11) * 0.25f;
    c[1] = (b00 + b01 - b10 - b11) * 0.25f;
    c[8] = (b00 - b01 + b10 - b11) * 0.25f;
    c[9] = (b00 - b01 - b10 + b11) * 0.25f;
    return;
  }
  for (int i = 0; i < 64; ++i) c[i] = 0.0f;
}


// Inverse transform of one block of raw strategy `raw` (TransformToPixels,
// pik/ac_strategy.cc:674-840). `coef` is the contiguous coefficient block:
// N x N row-major for DCT16/32 (the reference gathers it the same way,
// pik/ac_strategy.cc:804-831), 8 x 8 otherwise.
PIKGPU_HD void BlockInverse(int raw, const float* coef, float* pix,
                            int pstride) {
  if (raw == kDCT || raw == kDCTNoHF) {
    dct1d::Idct2D<8>(coef, 8, pix, pstride);
  } else if (raw == kDCT16) {
    dct1d::Idct2D<16>(coef, 16, pix, pstride);
  } else if (raw == kDCT32) {
    dct1d::Idct2D<32>(coef, 32, pix, pstride);
  } else if (raw == kDCT4X4 || raw == kDCT4X4NoHF) {
    const float b00 = coef[0], b01 = coef[1], b10 = coef[8], b11 = coef[9];
    float dcs[4];
    dcs[0] = b00 + b01 + b10 + b11;
    dcs[1] = b00 + b01 - b10 - b11;
    dcs[2] = b00 - b01 + b10 - b11;
    dcs[3] = b00 - b01 - b10 + b11;
    for (int y = 0; y < 2; ++y) {
      for (int x = 0; x < 2; ++x) {
        float blk[16];
        blk[0] = dcs[y * 2 + x];
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            if (ix == 0 && iy == 0) continue;
            blk[iy * 4 + ix] = coef[(y + iy * 2) * 8 + x + ix * 2];
          }
        }
        dct1d::Idct2D<4>(blk, 4, pix + y * 4 * pstride + x * 4, pstride);
      }
    }
  } else {
    DirectInverse8(raw, coef, pix, pstride);
  }
}

// Forward transform (TransformFromPixels, pik/ac_strategy.cc:531-672).
PIKGPU_HD void BlockForward(int raw, const float* pix, int pstride,
                            float* coef) {
  if (raw == kDCT || raw == kDCTNoHF) {
    dct1d::Dct2D<8>(pix, pstride, coef, 8);
  } else if (raw == kDCT16) {
    dct1d::Dct2D<16>(pix, pstride, coef, 16);
  } else if (raw == kDCT32) {
    dct1d::Dct2D<32>(pix, pstride, coef, 32);
  } else if (raw == kDCT4X4 || raw == kDCT4X4NoHF) {
    for (int y = 0; y < 2; ++y) {
      for (int x = 0; x < 2; ++x) {
        float blk[16];
        dct1d::Dct2D<4>(pix + y * 4 * pstride + x * 4, pstride, blk, 4);
        for (int iy = 0; iy < 4; ++iy) {
          for (int ix = 0; ix < 4; ++ix) {
            coef[(y + iy * 2) * 8 + x + ix * 2] = blk[iy * 4 + ix];
          }
        }
      }
    }
    const float b00 = coef[0], b01 = coef[1], b10 = coef[8], b11 = coef[9];
    coef[0] = (b00 + b01 + b10 + b11) * 0.25f;
    coef[1] = (b00 + b01 - b10 - b11) * 0.25f;
    coef[8] = (b00 - b01 + b10 - b11) * 0.25f;
    coef[9] = (b00 - b01 - b10 + b11) * 0.25f;
  } else {
    DirectForward8(raw, pix, pstride, coef);
  }
}

}  // namespace pikgpu
