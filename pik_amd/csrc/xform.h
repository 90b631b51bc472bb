// Per-strategy block transforms shared by the device kernels and the host
// self-check library: the direct 8x8 forms (DCT2X2, IDENTITY) and whole-block
// transposed-scaled transforms built on dct1d.h.
#pragma once

#include "common.h"
#include "dct1d.h"

namespace pikgpu {

// ---------------------------------------------------------------------------
// LINES: the diagonal-lines transform of an 8x8 block (pik/ac_strategy.cc:
// 380-530 for its tables and 1-D transforms, 545-568 forward, 680-724
// inverse). Pixels and coefficients are addressed row-major (y * 8 + x)
// here. Thirteen 45-degree lines of 3, 4, 6 and 8 pixels carry 1-D DCTs of
// their length into 60 coefficient slots; four pixel groups (a 2x2
// partition of the block along the lines) each carry their mean, and the
// four means a 2x2 Hadamard in slots 0, 8, 1 and 9; four pixels on no line
// are interpolated from their vertical neighbours. The 1-D transforms keep
// the reference's float / double evaluation: a product with a double
// constant is formed in double and rounded where the reference assigns it
// to a float.

// Line k: length kLinesLen[k], pixel positions kLinesPix[k][0..len), the
// slots of its coefficients kLinesCoef[k][0..len).
constexpr int kLinesCount = 13;
constexpr uint8_t kLinesLen[kLinesCount] = {3, 3, 3, 3, 4, 4, 4, 4,
                                            6, 6, 6, 6, 8};
constexpr uint8_t kLinesPix[kLinesCount][8] = {
    {48, 56, 57}, {40, 49, 58}, {5, 14, 23}, {6, 7, 15},
    {32, 41, 50, 59}, {24, 33, 51, 60}, {3, 12, 30, 39}, {4, 13, 22, 31},
    {16, 25, 34, 43, 52, 61}, {8, 17, 26, 44, 53, 62},
    {1, 10, 19, 37, 46, 55}, {2, 11, 20, 29, 38, 47},
    {0, 9, 18, 27, 36, 45, 54, 63}};
constexpr uint8_t kLinesCoef[kLinesCount][8] = {
    {48, 56, 57}, {40, 49, 58}, {5, 14, 23}, {6, 7, 15},
    {32, 41, 50, 59}, {24, 33, 42, 51}, {3, 12, 21, 30}, {4, 13, 22, 31},
    {16, 25, 34, 43, 52, 61}, {17, 26, 35, 44, 53, 62},
    {10, 19, 28, 37, 46, 55}, {2, 11, 20, 29, 38, 47},
    {18, 27, 36, 45, 54, 60, 39, 63}};
// The four mean groups (their sums run in this order) and the slots of the
// Hadamard of their means: group g's mean is Hadamard input g.
constexpr uint8_t kLinesGroupLen[4] = {21, 15, 15, 13};
constexpr uint8_t kLinesGroup[4][21] = {
    {0, 1, 9, 2, 10, 18, 11, 19, 27, 20, 28, 36, 29, 37, 45, 38, 46, 54, 47,
     55, 63},
    {3, 4, 12, 5, 13, 21, 6, 14, 22, 30, 7, 15, 23, 31, 39},
    {24, 32, 40, 48, 56, 33, 41, 49, 57, 42, 50, 58, 51, 59, 60},
    {8, 16, 17, 25, 26, 34, 35, 43, 44, 52, 53, 61, 62}};
constexpr uint8_t kLinesGroupSlot[4] = {0, 8, 1, 9};
// Interpolated pixels: (pixel, above, below).
constexpr uint8_t kLinesInterp[4][3] = {
    {21, 12, 30}, {28, 19, 37}, {35, 26, 44}, {42, 33, 51}};

PIKGPU_HD float LinesDbl(float v, double k) {
  return static_cast<float>(static_cast<double>(v) * k);
}

// In-place 1-D inverse of length n = 3, 4, 6 or 8.
PIKGPU_HD void LinesIdct(int n, float* v) {
  if (n == 3) {
    const float t0 = v[0], t1 = v[1] * 0.86602540f;
    const float t2a = v[2] * 0.5f, t2b = v[2];
    v[0] = t0 + t2a + t1;
    v[1] = t0 - t2b;
    v[2] = t0 + t2a - t1;
  } else if (n == 4) {
    const float c2_8 = 0.7071067811865475244f;
    const float i0 = v[0] * (0.5f * 2), i1 = v[1] * (0.382683432365089772f * 2);
    const float i2 = v[2] * (0.5f * 2), i3 = v[3] * (0.923879532511286756f * 2);
    const float t0 = i0 + i2, t1 = i0 - i2, t2 = i1 + i3, t3 = i1 - i3;
    const float t4 = t3 * c2_8;
    const float t5 = t2 + t4;
    v[0] = t0 + t5;
    v[1] = t1 + t4;
    v[2] = t1 - t4;
    v[3] = t0 - t5;
  } else if (n == 6) {
    const float a0 = v[0];
    const float a1 = LinesDbl(v[1], 0.96592583), b1 = LinesDbl(v[1], 0.70710678),
                c1 = LinesDbl(v[1], 0.25881905);
    const float a2 = LinesDbl(v[2], 0.86602540);
    const float a3 = LinesDbl(v[3], 0.70710678);
    const float a4 = LinesDbl(v[4], 0.5), b4 = v[4];
    const float a5 = LinesDbl(v[5], 0.25881905), b5 = LinesDbl(v[5], 0.70710678),
                c5 = LinesDbl(v[5], 0.96592583);
    v[0] = a0 + a1 + a3 + a4 + a5 + a2;
    v[1] = a0 + b1 - a3 - b4 - b5;
    v[2] = a0 + c1 - a3 + a4 + c5 - a2;
    v[3] = a0 - c1 + a3 + a4 - c5 - a2;
    v[4] = a0 - b1 + a3 - b4 + b5;
    v[5] = a0 - a1 - a3 + a4 - a5 + a2;
  } else {
    const float c1 = float(1.41421356237310), c2 = float(2.61312592975275);
    const float c3 = float(0.76536686473018), c4 = float(1.08239220029239);
    const double k = double(2.8284271247461903f);
    const float i0 = LinesDbl(v[0], 0.353553390593273762 * k);
    const float i1 = LinesDbl(v[1], 0.490392640201615225 * k);
    const float i2 = LinesDbl(v[2], 0.461939766255643378 * k);
    const float i3 = LinesDbl(v[3], 0.415734806151272619 * k);
    const float i4 = LinesDbl(v[4], 0.353553390593273762 * k);
    const float i5 = LinesDbl(v[5], 0.277785116509801112 * k);
    const float i6 = LinesDbl(v[6], 0.191341716182544886 * k);
    const float i7 = LinesDbl(v[7], 0.0975451610080641339 * k);
    const float t00 = i0 + i4, t01 = i0 - i4, t02 = i6 + i2, t03 = i6 - i2;
    const float t04 = i7 + i1, t05 = i7 - i1, t06 = i5 + i3, t07 = i5 - i3;
    const float t08 = t04 + t06, t09 = t04 - t06, t10 = t00 + t02;
    const float t11 = t00 - t02, t12 = t07 - t05;
    const float t13 = c3 * t12;
    const float t14 = c1 * t03 + t02;
    const float t15 = t01 - t14, t16 = t01 + t14;
    const float t17 = c2 * t05 + t13;
    const float t18 = c4 * t07 + t13;
    const float t19 = t08 + t17;
    const float t20 = c1 * t09 + t19;
    const float t21 = t18 - t20;
    v[0] = t10 + t08;
    v[1] = t15 - t19;
    v[2] = t16 + t20;
    v[3] = t11 + t21;
    v[4] = t11 - t21;
    v[5] = t16 - t20;
    v[6] = t15 + t19;
    v[7] = t10 - t08;
  }
}

// In-place 1-D forward of length n = 3, 4, 6 or 8.
PIKGPU_HD void LinesDct(int n, float* v) {
  if (n == 3) {
    const float i0 = v[0], i1 = v[1], i2 = v[2];
    v[0] = (i0 + i1 + i2) * (1 / 3.0f);
    v[1] = LinesDbl(i0 - i2, 0.57735029);
    v[2] = (i0 + i2) * (1 / 3.0f) - i1 * (2 / 3.0f);
  } else if (n == 4) {
    const float c2_8 = 1.414213562373095048f;
    const float t0 = v[0] + v[3], t1 = v[1] + v[2], t2 = v[0] - v[3];
    const float t3 = v[1] - v[2];
    const float t4 = t0 + t1, t5 = t0 - t1, t6 = t2 - t3, t7 = t3 * c2_8;
    const float t8 = t6 + t7, t9 = t6 - t7;
    v[0] = t4 * (0.5f / 2);
    v[1] = t8 * (0.653281482438188264f / 2);
    v[2] = t5 * (0.5f / 2);
    v[3] = t9 * (0.270598050073098492f / 2);
  } else if (n == 6) {
    const float i0 = v[0], i1 = v[1], i2 = v[2], i3 = v[3], i4 = v[4],
                i5 = v[5];
    const float t0 = LinesDbl(i1 - i4, 0.23570227);
    v[0] = (i0 + i1 + i2 + i3 + i4 + i5) * (1 / 6.0f);
    v[1] = static_cast<float>(double(i0 - i5) * 0.32197529 + double(t0) +
                              double(i2 - i3) * 0.08627302);
    v[2] = LinesDbl(i0 - i2 - i3 + i5, 0.28867514);
    v[3] = LinesDbl(i0 - i1 - i2 + i3 + i4 - i5, 0.23570227);
    v[4] = (i0 + i2 + i3 + i5) * (1 / 6.0f) - (i1 + i4) * (1 / 3.0f);
    v[5] = static_cast<float>(double(i0 - i5) * 0.08627302 - double(t0) +
                              double(i2 - i3) * 0.32197529);
  } else {
    const float c1 = 0.707106781186548f, c2 = 0.382683432365090f;
    const float c3 = 1.30656296487638f, c4 = 0.541196100146197f;
    const float t00 = v[0] + v[7], t01 = v[0] - v[7], t02 = v[3] + v[4];
    const float t03 = v[3] - v[4], t04 = v[2] + v[5], t05 = v[2] - v[5];
    const float t06 = v[1] + v[6], t07 = v[1] - v[6];
    const float t08 = t00 + t02, t09 = t00 - t02, t10 = t06 + t04;
    const float t11 = t06 - t04, t12 = t07 + t05, t13 = t01 + t07;
    const float t14 = t05 + t03, t15 = t11 + t09, t16 = t14 - t13;
    const float t17 = c1 * t15, t18 = c1 * t12, t19 = c2 * t16;
    const float t20 = t01 + t18, t21 = t01 - t18;
    const float t22 = c3 * t13 + t19;
    const float t23 = c4 * t14 + t19;
    const float k = 2.8284271247461903f;
    v[0] = (t08 + t10) * (0.353553390593273762f / k);
    v[1] = (t20 + t22) * (0.254897789552079584f / k);
    v[2] = (t09 + t17) * (0.270598050073098492f / k);
    v[3] = (t21 - t23) * (0.30067244346752264f / k);
    v[4] = (t08 - t10) * (0.353553390593273762f / k);
    v[5] = (t21 + t23) * (0.449988111568207852f / k);
    v[6] = (t09 - t17) * (0.653281482438188264f / k);
    v[7] = (t20 - t22) * (1.28145772387075309f / k);
  }
}

// Inverse LINES of one block: c holds the 64 coefficient slots, p the
// pixels (row stride pstride).
PIKGPU_HD void LinesInverse8(const float* c, float* p, int pstride) {
  float t[64];
  #pragma unroll
  for (int i = 0; i < 64; ++i) t[i] = 0.0f;
  #pragma unroll
  for (int l = 0; l < kLinesCount; ++l) {
    const int n = kLinesLen[l];
    float v[8];
    #pragma unroll
    for (int j = 0; j < n; ++j) v[j] = c[kLinesCoef[l][j]];
    LinesIdct(n, v);
    #pragma unroll
    for (int j = 0; j < n; ++j) t[kLinesPix[l][j]] = v[j];
  }
  const float h0 = c[0], h1 = c[8], h2 = c[1], h3 = c[9];
  float dc[4];
  dc[0] = h0 + h1 + h2 + h3;
  dc[1] = h0 - h1 + h2 - h3;
  dc[2] = h0 + h1 - h2 - h3;
  dc[3] = h0 - h1 - h2 + h3;
  #pragma unroll
  for (int g = 0; g < 4; ++g) {
    #pragma unroll
    for (int j = 0; j < kLinesGroupLen[g]; ++j) {
      t[kLinesGroup[g][j]] += dc[g];
    }
  }
  #pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint8_t* q = kLinesInterp[k];
    t[q[0]] = (t[q[1]] + t[q[2]]) * 0.5f;
  }
  #pragma unroll
  for (int y = 0; y < 8; ++y) {
    #pragma unroll
    for (int x = 0; x < 8; ++x) p[y * pstride + x] = t[y * 8 + x];
  }
}

// Forward LINES of one block (p: pixels, row stride pstride; c: 64 slots).
PIKGPU_HD void LinesForward8(const float* p, int pstride, float* c) {
  float t[64];
  #pragma unroll
  for (int y = 0; y < 8; ++y) {
    #pragma unroll
    for (int x = 0; x < 8; ++x) t[y * 8 + x] = p[y * pstride + x];
  }
  float dc[4];
  #pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int n = kLinesGroupLen[g];
    float sum = 0.0f;
    #pragma unroll
    for (int j = 0; j < n; ++j) sum += t[kLinesGroup[g][j]];
    sum /= float(n);
    #pragma unroll
    for (int j = 0; j < n; ++j) t[kLinesGroup[g][j]] -= sum;
    dc[g] = sum;
  }
  const float i0 = dc[0], i1 = dc[1], i2 = dc[2], i3 = dc[3];
  c[kLinesGroupSlot[0]] = (i0 + i1 + i2 + i3) * 0.25f;
  c[kLinesGroupSlot[1]] = (i0 - i1 + i2 - i3) * 0.25f;
  c[kLinesGroupSlot[2]] = (i0 + i1 - i2 - i3) * 0.25f;
  c[kLinesGroupSlot[3]] = (i0 - i1 - i2 + i3) * 0.25f;
  #pragma unroll
  for (int l = 0; l < kLinesCount; ++l) {
    const int n = kLinesLen[l];
    float v[8];
    #pragma unroll
    for (int j = 0; j < n; ++j) v[j] = t[kLinesPix[l][j]];
    LinesDct(n, v);
    #pragma unroll
    for (int j = 0; j < n; ++j) c[kLinesCoef[l][j]] = v[j];
  }
}

// Direct 8x8 inverse transforms: DCT2X2 pyramid (pik/ac_strategy.cc:
// 791-803, 330-356), IDENTITY (pik/ac_strategy.cc:725-761) and LINES
// (pik/ac_strategy.cc:680-724).
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
  if (raw == kLines) {
    LinesInverse8(c, p, pstride);
    return;
  }
  for (int y = 0; y < 8; ++y) {
    for (int x = 0; x < 8; ++x) p[y * pstride + x] = 0.0f;
  }
}

// Direct 8x8 forward transforms (pik/ac_strategy.cc:569-589, 614-618;
// LINES 545-568).
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
  if (raw == kLines) {
    LinesForward8(p, pstride, c);
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
