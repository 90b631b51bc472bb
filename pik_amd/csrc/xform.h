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

// ---------------------------------------------------------------------------
// Adjoints (transposes) of the inverse transforms, for the gradient of the
// reconstruction with respect to the coefficients.
//
// The 1-D transposed-scaled IDCT is pix(x) = sum_k F_k b_k(x) with mutually
// orthogonal basis rows b_k (b_k(x) = w_k cos((2x+1) k pi / 2N); w_0 = 1,
// w_k = 1/cos(k pi/16) for N = 8 and 2 cos(k pi/2N) otherwise). Its forward
// transform is the exact inverse, F_k = (1/D_k) sum_x pix(x) b_k(x) with
// D_k = sum_x b_k(x)^2, so the transpose is the forward transform with
// coefficient k scaled by D_k: D_0 = N, D_k = 4/cos^2(k pi/16) (N = 8),
// D_k = 2N cos^2(k pi/2N) (N = 4, 16, 32). In 2-D coefficient (i, j) is
// scaled by D_i D_j. DCT4X4: the same per 4x4 sub-block, and the four DCs
// pass through the DC Hadamard H (H H = 4 I), whose transpose is H = 4 H/4:
// scale 4 D_0^2 = 64. DCT2X2: each of the three Hadamard levels S = 8, 4, 2
// (region S x S) transposes to 4x its inverse level, so a coefficient gets
// 4 per level that covers it: 64 in the 2x2 corner, 16 in the rest of the
// 4x4 corner, 4 elsewhere. IDENTITY has its own adjoint (IdentityAdjoint8).
// Scale of coefficient `flat` (i * N + j of the N x N coefficient matrix in
// storage order; 0..63 for the 8x8 strategies) of a transform of raw
// strategy `raw` applied after the forward transform. `norms` holds D_k for
// N = 4, 8, 16, 32 at AdjNormOffset(N).
PIKGPU_HD float AdjointScale(const float* norms, int raw, int flat) {
  if (AcsIsDCT(raw)) {
    const int n = 8 * AcsCovered(raw);
    const float* d = norms + AdjNormOffset(n);
    return d[flat / n] * d[flat % n];
  }
  const int ky = flat >> 3, kx = flat & 7;
  if (AcsIsDCT4(raw)) {
    if (ky < 2 && kx < 2) return 64.0f;
    const float* d = norms + AdjNormOffset(4);
    return d[ky >> 1] * d[kx >> 1];
  }
  if (raw == kDCT2X2) {
    return (ky < 2 && kx < 2) ? 64.0f : ((ky < 4 && kx < 4) ? 16.0f : 4.0f);
  }
  return 1.0f;
}

// Transpose of DirectInverse8(kIdentity): every pixel of a 4x4 sub-block
// holds the sub-block's centre value, which is its Hadamard DC minus 1/16
// of the sum of its 15 non-DC coefficients; the other pixels add one
// coefficient each (the one stored at (1, 1) goes to pixel (0, 0)).
PIKGPU_HD void IdentityAdjoint8(const float* p, int pstride, float* c) {
  float dcs[4];
  for (int y = 0; y < 2; ++y) {
    for (int x = 0; x < 2; ++x) {
      float sum = 0.0f;
      for (int iy = 0; iy < 4; ++iy) {
        for (int ix = 0; ix < 4; ++ix) {
          sum += p[(4 * y + iy) * pstride + 4 * x + ix];
        }
      }
      dcs[y * 2 + x] = sum;
      const float share = sum * (1.0f / 16);
      for (int iy = 0; iy < 4; ++iy) {
        for (int ix = 0; ix < 4; ++ix) {
          if ((iy == 0 && ix == 0) || (iy == 1 && ix == 1)) continue;
          c[(y + 2 * iy) * 8 + x + 2 * ix] =
              p[(4 * y + iy) * pstride + 4 * x + ix] - share;
        }
      }
      c[(y + 2) * 8 + x + 2] = p[4 * y * pstride + 4 * x] - share;
    }
  }
  c[0] = dcs[0] + dcs[1] + dcs[2] + dcs[3];
  c[1] = dcs[0] + dcs[1] - dcs[2] - dcs[3];
  c[8] = dcs[0] - dcs[1] + dcs[2] - dcs[3];
  c[9] = dcs[0] - dcs[1] - dcs[2] + dcs[3];
}

// Transpose of BlockInverse (contiguous N x N in and out).
PIKGPU_HD void BlockAdjoint(int raw, const float* norms, const float* pix,
                            int pstride, float* coef) {
  if (raw == kIdentity) {
    IdentityAdjoint8(pix, pstride, coef);
    return;
  }
  BlockForward(raw, pix, pstride, coef);
  const int n = 8 * AcsCovered(raw);
  if (AcsIsDCT(raw) || AcsIsDCT4(raw) || raw == kDCT2X2) {
    for (int i = 0; i < n * n; ++i) coef[i] *= AdjointScale(norms, raw, i);
  }
}

}  // namespace pikgpu
