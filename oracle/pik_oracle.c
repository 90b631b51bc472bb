/*
 * pik_oracle.c -- CPU restatement of google/pik's decode-side reconstruction:
 * DequantDC + AdaptiveDCReconstruction, per-group DC cache, lowest- and
 * low-frequency predictions, dequantisation with chroma-from-luma, HF
 * prediction, the variable-size IDCT, the Gaborish guide, the edge-
 * preserving filter and the AR clamp, dot patches, the final Gaborish and
 * OpsinToLinear.
 *
 * TEST INFRASTRUCTURE (the oracle). Only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark use it, as the checker; the product
 * never links it.
 *
 * It is written frame at a time, single-threaded, in plain loops, and keeps
 * the floating-point operation order of the reference's scalar build (built
 * with -ffp-contract=off) so that it agrees with the reference bit for bit.
 * Every function cites the reference code it restates. Parity: each stage is
 * pinned bit for bit against the reference compiled from its own sources
 * (oracle/_ref) and against the golden fixtures under tests/golden
 * (tests/test_oracle.py); the DCT scale tables and butterflies are pinned
 * against their closed forms as well.
 */
#include "pik_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static char g_err[256];

static int check_frame(const pikgpu_frame* f);

static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

const char* pikoracle_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* Geometry helpers                                                          */

enum {
  kDCT = 0, kIdentity = 1, kDCT2X2 = 2, kDCT4X4 = 3, kDCT16 = 4, kDCT32 = 5,
  kLines = 6, kDCTNoHF = 7, kDCT4X4NoHF = 8
};
enum { kGroupBlocks = 64, kBorder = 6, kMinSigma = 16, kMaxSigma = 672 };

/* Mirror, pik/image_ops.h:217-227: mirrors outside [0, n), repeating the
 * edge pixel. */
static long mirror(long x, long n) {
  while (x < 0 || x >= n) x = x < 0 ? -x - 1 : 2 * n - 1 - x;
  return x;
}

static size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }

/* FrameHeader::kGrayscaleOpt (pik/headers.h:185-196) from either of the
 * frame's two spellings. */
static int is_gray(const pikgpu_frame* f) {
  return f->grayscale != 0 || (f->flags & PIKGPU_FLAG_GRAYSCALE_OPT) != 0;
}

/* AcStrategy properties, pik/ac_strategy.h:41-170. */
static int covered(int raw) {
  return raw == kDCT32 ? 4 : (raw == kDCT16 ? 2 : 1);
}
static int quant_kind(int raw) {
  if (raw == kDCTNoHF) return 0;
  if (raw == kDCT4X4NoHF) return kDCT4X4;
  return raw;
}
static int predicts_hf(int raw) {
  return raw != kDCT2X2 && raw != kIdentity && raw != kDCTNoHF &&
         raw != kDCT4X4NoHF && raw != kLines;
}
static float ar_quant_scale(int raw) {
  if (raw == kDCT32) return (float)1.2282996852328099;
  if (raw == kDCT16) return (float)1.1423171621463439;
  if (raw == kDCT4X4 || raw == kDCT4X4NoHF) return 1.2f;
  if (raw == kDCT2X2) return (float)1.0098134203870499;
  return 1.0f;
}

/* ------------------------------------------------------------------------ */
/* Scale tables of the transposed-scaled transforms (pik/dct.h:25-260).     */
/* The reference lists them as literals produced by a documented formula:   */
/* iscale(u) = sqrt(2/N) eps(u) cos(u pi / 2N), eps(0) = sqrt(1/2), and     */
/* scale(u) = 1 / (N iscale(u)); the AAN 8-point DCT uses DCTScales = scale */
/* and IDCTScales = iscale, the recursive algorithm of the other sizes the  */
/* other way round. L1Norm(i) is the product of cos(pi/8), cos(pi/16), ...  */
/* over log2(N) - 1 - (trailing zero bits of i) factors (1 for i = 0).      */

static float g_dct_scale[33][32], g_idct_scale[33][32];
static float g_l1[33][32], g_l1inv[33][32];
static float g_fwd_c[33][64], g_inv_c[33][64];
static int g_mul_table[kMaxSigma + 1];
static int g_init;

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

static void init_weight_table(void);

static void init_tables(void) {
  if (g_init) return;
  for (int n = 2; n <= 32; n *= 2) {
    for (int u = 0; u < n; ++u) {
      const double eps = u == 0 ? sqrt(0.5) : 1.0;
      const double iscale = sqrt(2.0 / n) * eps * cos(u * M_PI / (2.0 * n));
      const double scale = 1.0 / (n * iscale);
      g_dct_scale[n][u] = (float)(n == 8 ? scale : iscale);
      g_idct_scale[n][u] = (float)(n == 8 ? iscale : scale);
      double l1 = 1.0;
      if (u != 0) {
        int tz = 0;
        while (((u >> tz) & 1) == 0) ++tz;
        const int factors = ilog2(n) - 1 - tz;
        for (int f = 0; f < factors; ++f) l1 *= cos(M_PI / (8 << f));
      }
      g_l1[n][u] = (float)l1;
      g_l1inv[n][u] = (float)(1.0 / l1);
    }
    /* Butterfly constants: 2 cos(J pi / 2n) (forward) and
     * 0.5 / cos(J pi / 2n) (inverse), pik/dct_simd_any.h. */
    for (int j = 1; j < n; ++j) {
      g_fwd_c[n][j] = (float)(2.0 * cos(j * M_PI / (2.0 * n)));
      g_inv_c[n][j] = (float)(0.5 / cos(j * M_PI / (2.0 * n)));
    }
  }
  init_weight_table();
  g_init = 1;
}

/* DCTTotalScale / DCTInvTotalScale, pik/ac_strategy.cc:258-270 (float). */
static float total_scale(int n, int x, int y) {
  return (float)n * g_dct_scale[n][x] * g_dct_scale[n][y] * g_l1inv[n][x] *
         g_l1inv[n][y];
}
static float inv_total_scale(int n, int x, int y) {
  return (float)n * g_idct_scale[n][x] * g_idct_scale[n][y] * g_l1[n][x] *
         g_l1[n][y];
}

/* ------------------------------------------------------------------------ */
/* One-dimensional butterflies (pik/dct.h:25-40, pik/dct_simd_8.h,          */
/* pik/dct_simd_any.h): the Arai-Agui-Nakajima flow graph for standalone    */
/* 8-point transforms, the recursive algorithm of Vashkevich and Petrovsky  */
/* for 4, 16 and 32 points (and the 8-point halves inside them).            */

/* Forward, odd part: a node of m inputs at angle J (units of pi/2n). */
static void vp_fwd_odd(int n, int m, const float* u, int J, float* y) {
  float p[16], q[16];
  const float c = g_fwd_c[n][J];
  for (int k = 0; k < m / 2; ++k) {
    const float a = u[k] - u[m - 1 - k];
    const float b = u[m / 2 + k] * c;
    p[k] = a + b;
    q[k] = a - b;
  }
  if (m == 2) {
    y[J / 2] = p[0];
    y[n - J / 2] = q[0];
    return;
  }
  vp_fwd_odd(n, m / 2, p, J / 2, y);
  vp_fwd_odd(n, m / 2, q, (2 * n - J) / 2, y);
}

static void vp_dct(int n, const float* x, float* y) {
  if (n == 2) {
    y[0] = x[0] + x[1];
    y[1] = x[0] - x[1];
    return;
  }
  float s[16], d[16], e[16];
  for (int k = 0; k < n / 2; ++k) {
    s[k] = x[k] + x[n - 1 - k];
    d[k] = x[k] - x[n - 1 - k];
  }
  vp_dct(n / 2, s, e);
  for (int k = 0; k < n / 2; ++k) y[2 * k] = e[k];
  vp_fwd_odd(n, n / 2, d, n / 2, y);
}

/* Inverse, odd part: merges the groups of angles J/2 and (2n - J)/2. */
static void vp_inv_odd(int n, int size, const float* x, int J, float* out) {
  const float c = g_inv_c[n][J];
  if (size == 2) {
    const int j = J / 2;
    const float s = x[j] + x[n - j];
    const float d = (x[j] - x[n - j]) * c;
    out[0] = s + d;
    out[1] = d;
    return;
  }
  float g[16], h[16], s[16], d[16];
  vp_inv_odd(n, size / 2, x, J / 2, g);
  vp_inv_odd(n, size / 2, x, (2 * n - J) / 2, h);
  for (int k = 0; k < size / 2; ++k) {
    s[k] = g[k] + h[k];
    d[k] = (g[k] - h[k]) * c;
  }
  for (int k = 0; k < size / 2; ++k) {
    out[k] = s[k] + d[size / 2 - 1 - k];
    out[size / 2 + k] = d[k];
  }
}

static void vp_idct(int n, const float* x, float* y) {
  if (n == 2) {
    y[0] = x[0] + x[1];
    y[1] = x[0] - x[1];
    return;
  }
  float ev[16], e[16], o[16];
  for (int k = 0; k < n / 2; ++k) ev[k] = x[2 * k];
  vp_idct(n / 2, ev, e);
  vp_inv_odd(n, n / 2, x, n / 2, o);
  for (int k = 0; k < n / 2; ++k) {
    y[k] = e[k] + o[k];
    y[n - 1 - k] = e[k] - o[k];
  }
}

static void aan_dct8(const float* i, float* o) {
  const float c1 = 0.707106781186548f;
  const float c2 = 0.382683432365090f;
  const float c3 = 1.30656296487638f;
  const float c4 = 0.541196100146197f;
  const float t00 = i[0] + i[7], t01 = i[0] - i[7];
  const float t02 = i[3] + i[4], t03 = i[3] - i[4];
  const float t04 = i[2] + i[5], t05 = i[2] - i[5];
  const float t06 = i[1] + i[6], t07 = i[1] - i[6];
  const float t08 = t00 + t02, t09 = t00 - t02;
  const float t10 = t06 + t04, t11 = t06 - t04;
  const float t12 = t07 + t05, t13 = t01 + t07, t14 = t05 + t03;
  const float t15 = t11 + t09;
  const float t16 = t14 - t13;
  const float t17 = c1 * t15;
  const float t18 = c1 * t12;
  const float t19 = c2 * t16;
  const float t20 = t01 + t18, t21 = t01 - t18;
  const float t22 = c3 * t13 + t19;
  const float t23 = c4 * t14 + t19;
  o[0] = t08 + t10;
  o[1] = t20 + t22;
  o[2] = t09 + t17;
  o[3] = t21 - t23;
  o[4] = t08 - t10;
  o[5] = t21 + t23;
  o[6] = t09 - t17;
  o[7] = t20 - t22;
}

static void aan_idct8(const float* i, float* o) {
  const float c1 = 1.41421356237310f;
  const float c2 = 2.61312592975275f;
  const float c3 = 0.76536686473018f;
  const float c4 = 1.08239220029239f;
  const float t00 = i[0] + i[4], t01 = i[0] - i[4];
  const float t02 = i[6] + i[2], t03 = i[6] - i[2];
  const float t04 = i[7] + i[1], t05 = i[7] - i[1];
  const float t06 = i[5] + i[3], t07 = i[5] - i[3];
  const float t08 = t04 + t06, t09 = t04 - t06;
  const float t10 = t00 + t02, t11 = t00 - t02;
  const float t12 = (t07 - t05) * c3;
  const float t13 = c1 * t03 + t02;
  const float t14 = t01 - t13, t15 = t01 + t13;
  const float t16 = c2 * t05 + t12;
  const float t17 = c4 * t07 + t12;
  const float t18 = t08 + t16;
  const float t19 = c1 * t09 + t18;
  const float t20 = t17 - t19;
  o[0] = t10 + t08;
  o[1] = t14 - t18;
  o[2] = t15 + t19;
  o[3] = t11 + t20;
  o[4] = t11 - t20;
  o[5] = t15 - t19;
  o[6] = t14 + t18;
  o[7] = t10 - t08;
}

static void dct_1d(int n, const float* x, float* y) {
  if (n == 8) aan_dct8(x, y); else vp_dct(n, x, y);
}
static void idct_1d(int n, const float* x, float* y) {
  if (n == 8) aan_idct8(x, y); else vp_idct(n, x, y);
}

/* ComputeTransposedScaledIDCT<n> (pik/dct.h:444-507): IDCT of the columns
 * of the coefficient matrix, transpose, IDCT of the columns. n = 2 is the
 * explicit 2x2 form. coef is row-major n x n, the row index being the
 * horizontal frequency. */
static void idct_2d(int n, const float* coef, size_t cs, float* pix,
                    size_t ps) {
  if (n == 2) {
    const float a00 = coef[0], a01 = coef[1], a10 = coef[cs],
                a11 = coef[cs + 1];
    pix[0] = a00 + a01 + a10 + a11;
    pix[1] = a00 + a01 - a10 - a11;
    pix[ps] = a00 - a01 + a10 - a11;
    pix[ps + 1] = a00 - a01 - a10 + a11;
    return;
  }
  float t[32 * 32], v[32], w[32];
  for (int c = 0; c < n; ++c) {
    for (int r = 0; r < n; ++r) v[r] = coef[r * cs + c];
    idct_1d(n, v, w);
    for (int x = 0; x < n; ++x) t[c * n + x] = w[x];
  }
  for (int x = 0; x < n; ++x) {
    for (int c = 0; c < n; ++c) v[c] = t[c * n + x];
    idct_1d(n, v, w);
    for (int y = 0; y < n; ++y) pix[y * ps + x] = w[y];
  }
}

/* ComputeTransposedScaledDCT<n> followed by ScaleToBlock's 1/n^2. */
static void dct_2d(int n, const float* pix, size_t ps, float* coef,
                   size_t cs) {
  if (n == 2) {
    const float a00 = pix[0], a01 = pix[1], a10 = pix[ps], a11 = pix[ps + 1];
    coef[0] = (a00 + a01 + a10 + a11) * 0.25f;
    coef[1] = (a00 + a01 - a10 - a11) * 0.25f;
    coef[cs] = (a00 - a01 + a10 - a11) * 0.25f;
    coef[cs + 1] = (a00 - a01 - a10 + a11) * 0.25f;
    return;
  }
  float t[32 * 32], v[32], w[32];
  for (int x = 0; x < n; ++x) {
    for (int y = 0; y < n; ++y) v[y] = pix[y * ps + x];
    dct_1d(n, v, w);
    for (int k = 0; k < n; ++k) t[x * n + k] = w[k];
  }
  const float scale = 1.0f / (float)(n * n);
  for (int k = 0; k < n; ++k) {
    for (int x = 0; x < n; ++x) v[x] = t[x * n + k];
    dct_1d(n, v, w);
    for (int j = 0; j < n; ++j) coef[j * cs + k] = w[j] * scale;
  }
}

/* ------------------------------------------------------------------------ */
/* Block transforms, pik/ac_strategy.cc:531-840. Coefficients are in the    */
/* frame-wide storage layout: a transform covering s x s blocks keeps its   */
/* coefficients in s storage rows of 64 s floats (`cs` apart), which is the */
/* row-major (8 s) x (8 s) coefficient matrix cut into s pieces.            */

static void hadamard_dc(const float* c, float* dcs) {
  const float b00 = c[0], b01 = c[1], b10 = c[8], b11 = c[9];
  dcs[0] = b00 + b01 + b10 + b11;
  dcs[1] = b00 + b01 - b10 - b11;
  dcs[2] = b00 - b01 + b10 - b11;
  dcs[3] = b00 - b01 - b10 + b11;
}

static void quarter_hadamard_in_place(float* c) {
  const float b00 = c[0], b01 = c[1], b10 = c[8], b11 = c[9];
  c[0] = (b00 + b01 + b10 + b11) * 0.25f;
  c[1] = (b00 + b01 - b10 - b11) * 0.25f;
  c[8] = (b00 - b01 + b10 - b11) * 0.25f;
  c[9] = (b00 - b01 - b10 + b11) * 0.25f;
}

/* DCT2TopBlock<S>, pik/ac_strategy.cc:301-328. */
static void dct2_top(int S, const float* block, size_t stride, float* out) {
  float temp[64];
  const int h = S / 2;
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < h; ++x) {
      const float c00 = block[y * 2 * stride + x * 2];
      const float c01 = block[y * 2 * stride + x * 2 + 1];
      const float c10 = block[(y * 2 + 1) * stride + x * 2];
      const float c11 = block[(y * 2 + 1) * stride + x * 2 + 1];
      float r00 = c00 + c01 + c10 + c11;
      float r01 = c00 + c01 - c10 - c11;
      float r10 = c00 - c01 + c10 - c11;
      float r11 = c00 - c01 - c10 + c11;
      r00 *= 0.25f;
      r01 *= 0.25f;
      r10 *= 0.25f;
      r11 *= 0.25f;
      temp[y * 8 + x] = r00;
      temp[y * 8 + h + x] = r01;
      temp[(y + h) * 8 + x] = r10;
      temp[(y + h) * 8 + h + x] = r11;
    }
  }
  for (int y = 0; y < S; ++y) {
    for (int x = 0; x < S; ++x) out[y * 8 + x] = temp[y * 8 + x];
  }
}

/* IDCT2TopBlock<S>, pik/ac_strategy.cc:330-357. */
static void idct2_top(int S, const float* block, size_t stride_out,
                      float* out) {
  float temp[64];
  const int h = S / 2;
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < h; ++x) {
      const float c00 = block[y * 8 + x];
      const float c01 = block[y * 8 + h + x];
      const float c10 = block[(y + h) * 8 + x];
      const float c11 = block[(y + h) * 8 + h + x];
      temp[y * 2 * 8 + x * 2] = c00 + c01 + c10 + c11;
      temp[y * 2 * 8 + x * 2 + 1] = c00 + c01 - c10 - c11;
      temp[(y * 2 + 1) * 8 + x * 2] = c00 - c01 + c10 - c11;
      temp[(y * 2 + 1) * 8 + x * 2 + 1] = c00 - c01 - c10 + c11;
    }
  }
  for (int y = 0; y < S; ++y) {
    for (int x = 0; x < S; ++x) out[y * stride_out + x] = temp[y * 8 + x];
  }
}

/* LINES, the diagonal-lines transform (pik/ac_strategy.cc:380-530 tables
 * and 1-D transforms, 545-568 forward, 680-724 inverse), with pixels and
 * coefficient slots numbered row-major (y * 8 + x): the reference numbers
 * both column-major through its P1 / C1 macros, so each of its indices i is
 * (i & 7) * 8 + (i >> 3) here. */
static const int kLn[13] = {3, 3, 3, 3, 4, 4, 4, 4, 6, 6, 6, 6, 8};
static const int kLpix[13][8] = {
    {48, 56, 57}, {40, 49, 58}, {5, 14, 23}, {6, 7, 15},
    {32, 41, 50, 59}, {24, 33, 51, 60}, {3, 12, 30, 39}, {4, 13, 22, 31},
    {16, 25, 34, 43, 52, 61}, {8, 17, 26, 44, 53, 62},
    {1, 10, 19, 37, 46, 55}, {2, 11, 20, 29, 38, 47},
    {0, 9, 18, 27, 36, 45, 54, 63}};
static const int kLcoef[13][8] = {
    {48, 56, 57}, {40, 49, 58}, {5, 14, 23}, {6, 7, 15},
    {32, 41, 50, 59}, {24, 33, 42, 51}, {3, 12, 21, 30}, {4, 13, 22, 31},
    {16, 25, 34, 43, 52, 61}, {17, 26, 35, 44, 53, 62},
    {10, 19, 28, 37, 46, 55}, {2, 11, 20, 29, 38, 47},
    {18, 27, 36, 45, 54, 60, 39, 63}};
static const int kLgn[4] = {21, 15, 15, 13};
static const int kLg[4][21] = {
    {0, 1, 9, 2, 10, 18, 11, 19, 27, 20, 28, 36, 29, 37, 45, 38, 46, 54, 47,
     55, 63},
    {3, 4, 12, 5, 13, 21, 6, 14, 22, 30, 7, 15, 23, 31, 39},
    {24, 32, 40, 48, 56, 33, 41, 49, 57, 42, 50, 58, 51, 59, 60},
    {8, 16, 17, 25, 26, 34, 35, 43, 44, 52, 53, 61, 62}};
static const int kLgslot[4] = {0, 8, 1, 9};

/* float * double constant, rounded to float (the reference assigns such a
 * product to a float). */
static float mul_d(float v, double k) { return (float)((double)v * k); }

static void lines_idct(int n, float* v) {
  if (n == 3) {
    float t0 = v[0], t1 = v[1] * 0.86602540f, t2a = v[2] * 0.5f, t2b = v[2];
    v[0] = t0 + t2a + t1;
    v[1] = t0 - t2b;
    v[2] = t0 + t2a - t1;
  } else if (n == 4) {
    float i0 = v[0] * (0.5f * 2), i1 = v[1] * (0.382683432365089772f * 2);
    float i2 = v[2] * (0.5f * 2), i3 = v[3] * (0.923879532511286756f * 2);
    float t0 = i0 + i2, t1 = i0 - i2, t2 = i1 + i3, t3 = i1 - i3;
    float t4 = t3 * 0.7071067811865475244f, t5 = t2 + t4;
    v[0] = t0 + t5;
    v[1] = t1 + t4;
    v[2] = t1 - t4;
    v[3] = t0 - t5;
  } else if (n == 6) {
    float a0 = v[0], a1 = mul_d(v[1], 0.96592583), b1 = mul_d(v[1], 0.70710678);
    float c1 = mul_d(v[1], 0.25881905), a2 = mul_d(v[2], 0.86602540);
    float a3 = mul_d(v[3], 0.70710678), a4 = mul_d(v[4], 0.5), b4 = v[4];
    float a5 = mul_d(v[5], 0.25881905), b5 = mul_d(v[5], 0.70710678);
    float c5 = mul_d(v[5], 0.96592583);
    v[0] = a0 + a1 + a3 + a4 + a5 + a2;
    v[1] = a0 + b1 - a3 - b4 - b5;
    v[2] = a0 + c1 - a3 + a4 + c5 - a2;
    v[3] = a0 - c1 + a3 + a4 - c5 - a2;
    v[4] = a0 - b1 + a3 - b4 + b5;
    v[5] = a0 - a1 - a3 + a4 - a5 + a2;
  } else {
    const double k = (double)2.8284271247461903f;
    float c1 = (float)1.41421356237310, c2 = (float)2.61312592975275;
    float c3 = (float)0.76536686473018, c4 = (float)1.08239220029239;
    float i0 = mul_d(v[0], 0.353553390593273762 * k);
    float i1 = mul_d(v[1], 0.490392640201615225 * k);
    float i2 = mul_d(v[2], 0.461939766255643378 * k);
    float i3 = mul_d(v[3], 0.415734806151272619 * k);
    float i4 = mul_d(v[4], 0.353553390593273762 * k);
    float i5 = mul_d(v[5], 0.277785116509801112 * k);
    float i6 = mul_d(v[6], 0.191341716182544886 * k);
    float i7 = mul_d(v[7], 0.0975451610080641339 * k);
    float t00 = i0 + i4, t01 = i0 - i4, t02 = i6 + i2, t03 = i6 - i2;
    float t04 = i7 + i1, t05 = i7 - i1, t06 = i5 + i3, t07 = i5 - i3;
    float t08 = t04 + t06, t09 = t04 - t06, t10 = t00 + t02, t11 = t00 - t02;
    float t12 = t07 - t05, t13 = c3 * t12, t14 = c1 * t03 + t02;
    float t15 = t01 - t14, t16 = t01 + t14, t17 = c2 * t05 + t13;
    float t18 = c4 * t07 + t13, t19 = t08 + t17, t20 = c1 * t09 + t19;
    float t21 = t18 - t20;
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

static void lines_dct(int n, float* v) {
  if (n == 3) {
    float i0 = v[0], i1 = v[1], i2 = v[2];
    v[0] = (i0 + i1 + i2) * (1 / 3.0f);
    v[1] = mul_d(i0 - i2, 0.57735029);
    v[2] = (i0 + i2) * (1 / 3.0f) - i1 * (2 / 3.0f);
  } else if (n == 4) {
    float t0 = v[0] + v[3], t1 = v[1] + v[2], t2 = v[0] - v[3];
    float t3 = v[1] - v[2];
    float t4 = t0 + t1, t5 = t0 - t1, t6 = t2 - t3;
    float t7 = t3 * 1.414213562373095048f;
    float t8 = t6 + t7, t9 = t6 - t7;
    v[0] = t4 * (0.5f / 2);
    v[1] = t8 * (0.653281482438188264f / 2);
    v[2] = t5 * (0.5f / 2);
    v[3] = t9 * (0.270598050073098492f / 2);
  } else if (n == 6) {
    float i0 = v[0], i1 = v[1], i2 = v[2], i3 = v[3], i4 = v[4], i5 = v[5];
    float t0 = mul_d(i1 - i4, 0.23570227);
    v[0] = (i0 + i1 + i2 + i3 + i4 + i5) * (1 / 6.0f);
    v[1] = (float)((double)(i0 - i5) * 0.32197529 + (double)t0 +
                   (double)(i2 - i3) * 0.08627302);
    v[2] = mul_d(i0 - i2 - i3 + i5, 0.28867514);
    v[3] = mul_d(i0 - i1 - i2 + i3 + i4 - i5, 0.23570227);
    v[4] = (i0 + i2 + i3 + i5) * (1 / 6.0f) - (i1 + i4) * (1 / 3.0f);
    v[5] = (float)((double)(i0 - i5) * 0.08627302 - (double)t0 +
                   (double)(i2 - i3) * 0.32197529);
  } else {
    const float c1 = 0.707106781186548f, c2 = 0.382683432365090f;
    const float c3 = 1.30656296487638f, c4 = 0.541196100146197f;
    const float k = 2.8284271247461903f;
    float t00 = v[0] + v[7], t01 = v[0] - v[7], t02 = v[3] + v[4];
    float t03 = v[3] - v[4], t04 = v[2] + v[5], t05 = v[2] - v[5];
    float t06 = v[1] + v[6], t07 = v[1] - v[6];
    float t08 = t00 + t02, t09 = t00 - t02, t10 = t06 + t04, t11 = t06 - t04;
    float t12 = t07 + t05, t13 = t01 + t07, t14 = t05 + t03;
    float t15 = t11 + t09, t16 = t14 - t13;
    float t17 = c1 * t15, t18 = c1 * t12, t19 = c2 * t16;
    float t20 = t01 + t18, t21 = t01 - t18;
    float t22 = c3 * t13 + t19, t23 = c4 * t14 + t19;
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

static void lines_to_pixels(const float* coef, float* pix, size_t ps) {
  float t[64], v[8], h[4];
  memset(t, 0, sizeof(t));
  for (int l = 0; l < 13; ++l) {
    for (int j = 0; j < kLn[l]; ++j) v[j] = coef[kLcoef[l][j]];
    lines_idct(kLn[l], v);
    for (int j = 0; j < kLn[l]; ++j) t[kLpix[l][j]] = v[j];
  }
  {
    const float i0 = coef[0], i1 = coef[8], i2 = coef[1], i3 = coef[9];
    h[0] = i0 + i1 + i2 + i3;
    h[1] = i0 - i1 + i2 - i3;
    h[2] = i0 + i1 - i2 - i3;
    h[3] = i0 - i1 - i2 + i3;
  }
  for (int g = 0; g < 4; ++g) {
    for (int j = 0; j < kLgn[g]; ++j) t[kLg[g][j]] += h[g];
  }
  t[21] = (t[12] + t[30]) * 0.5f;
  t[28] = (t[19] + t[37]) * 0.5f;
  t[35] = (t[26] + t[44]) * 0.5f;
  t[42] = (t[33] + t[51]) * 0.5f;
  for (int y = 0; y < 8; ++y) {
    for (int x = 0; x < 8; ++x) pix[y * ps + x] = t[y * 8 + x];
  }
}

static void lines_from_pixels(const float* pix, size_t ps, float* coef) {
  float t[64], v[8], m[4];
  for (int y = 0; y < 8; ++y) {
    for (int x = 0; x < 8; ++x) t[y * 8 + x] = pix[y * ps + x];
  }
  for (int g = 0; g < 4; ++g) {
    float sum = 0;
    for (int j = 0; j < kLgn[g]; ++j) sum += t[kLg[g][j]];
    sum /= (float)kLgn[g];
    for (int j = 0; j < kLgn[g]; ++j) t[kLg[g][j]] -= sum;
    m[g] = sum;
  }
  coef[kLgslot[0]] = (m[0] + m[1] + m[2] + m[3]) * 0.25f;
  coef[kLgslot[1]] = (m[0] - m[1] + m[2] - m[3]) * 0.25f;
  coef[kLgslot[2]] = (m[0] + m[1] - m[2] - m[3]) * 0.25f;
  coef[kLgslot[3]] = (m[0] - m[1] - m[2] + m[3]) * 0.25f;
  for (int l = 0; l < 13; ++l) {
    for (int j = 0; j < kLn[l]; ++j) v[j] = t[kLpix[l][j]];
    lines_dct(kLn[l], v);
    for (int j = 0; j < kLn[l]; ++j) coef[kLcoef[l][j]] = v[j];
  }
}

/* AcStrategy::TransformToPixels, pik/ac_strategy.cc:674-840. */
static void to_pixels(int raw, const float* coef, size_t cs, float* pix,
                      size_t ps) {
  switch (raw) {
    case kDCT:
    case kDCTNoHF:
      idct_2d(8, coef, 8, pix, ps);
      break;
    case kDCT16:
    case kDCT32: {
      const int s = covered(raw), n = 8 * s;
      float in[32 * 32];
      for (int i = 0; i < s; ++i) {
        memcpy(in + 64 * s * i, coef + cs * i, sizeof(float) * 64 * s);
      }
      idct_2d(n, in, n, pix, ps);
      break;
    }
    case kDCT4X4:
    case kDCT4X4NoHF: {
      float dcs[4];
      hadamard_dc(coef, dcs);
      for (int y = 0; y < 2; ++y) {
        for (int x = 0; x < 2; ++x) {
          float block[16];
          block[0] = dcs[y * 2 + x];
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              if (ix == 0 && iy == 0) continue;
              block[iy * 4 + ix] = coef[(y + iy * 2) * 8 + x + ix * 2];
            }
          }
          idct_2d(4, block, 4, pix + y * 4 * ps + x * 4, ps);
        }
      }
      break;
    }
    case kDCT2X2: {
      float c[64];
      memcpy(c, coef, sizeof(c));
      idct2_top(2, c, 8, c);
      idct2_top(4, c, 8, c);
      idct2_top(8, c, 8, c);
      for (int y = 0; y < 8; ++y) {
        for (int x = 0; x < 8; ++x) pix[y * ps + x] = c[y * 8 + x];
      }
      break;
    }
    case kIdentity: {
      float dcs[4];
      hadamard_dc(coef, dcs);
      for (int y = 0; y < 2; ++y) {
        for (int x = 0; x < 2; ++x) {
          const float block_dc = dcs[y * 2 + x];
          float residual_sum = 0;
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              if (ix == 0 && iy == 0) continue;
              residual_sum += coef[(y + iy * 2) * 8 + x + ix * 2];
            }
          }
          float* p11 = pix + (4 * y + 1) * ps + 4 * x + 1;
          *p11 = block_dc - residual_sum * (1.0f / 16);
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              if (ix == 1 && iy == 1) continue;
              pix[(y * 4 + iy) * ps + x * 4 + ix] =
                  coef[(y + iy * 2) * 8 + x + ix * 2] + *p11;
            }
          }
          pix[y * 4 * ps + x * 4] = coef[(y + 2) * 8 + x + 2] + *p11;
        }
      }
      break;
    }
    case kLines:
      lines_to_pixels(coef, pix, ps);
      break;
    default:
      break;
  }
}

/* AcStrategy::TransformFromPixels, pik/ac_strategy.cc:531-672. */
static void from_pixels(int raw, const float* pix, size_t ps, float* coef,
                        size_t cs) {
  switch (raw) {
    case kDCT:
    case kDCTNoHF:
      dct_2d(8, pix, ps, coef, 8);
      break;
    case kDCT16:
    case kDCT32: {
      const int s = covered(raw), n = 8 * s;
      float out[32 * 32];
      dct_2d(n, pix, ps, out, n);
      for (int i = 0; i < s; ++i) {
        memcpy(coef + cs * i, out + 64 * s * i, sizeof(float) * 64 * s);
      }
      break;
    }
    case kDCT4X4:
    case kDCT4X4NoHF: {
      for (int y = 0; y < 2; ++y) {
        for (int x = 0; x < 2; ++x) {
          float block[16];
          dct_2d(4, pix + y * 4 * ps + x * 4, ps, block, 4);
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              coef[(y + iy * 2) * 8 + x + ix * 2] = block[iy * 4 + ix];
            }
          }
        }
      }
      quarter_hadamard_in_place(coef);
      break;
    }
    case kDCT2X2:
      dct2_top(8, pix, ps, coef);
      dct2_top(4, coef, 8, coef);
      dct2_top(2, coef, 8, coef);
      break;
    case kIdentity: {
      for (int y = 0; y < 2; ++y) {
        for (int x = 0; x < 2; ++x) {
          float block_dc = 0;
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              block_dc += pix[(y * 4 + iy) * ps + x * 4 + ix];
            }
          }
          block_dc *= 1.0f / 16;
          for (int iy = 0; iy < 4; ++iy) {
            for (int ix = 0; ix < 4; ++ix) {
              if (ix == 1 && iy == 1) continue;
              coef[(y + iy * 2) * 8 + x + ix * 2] =
                  pix[(y * 4 + iy) * ps + x * 4 + ix] -
                  pix[(y * 4 + 1) * ps + x * 4 + 1];
            }
          }
          coef[(y + 2) * 8 + x + 2] = coef[y * 8 + x];
          coef[y * 8 + x] = block_dc;
        }
      }
      quarter_hadamard_in_place(coef);
      break;
    }
    case kLines:
      lines_from_pixels(pix, ps, coef);
      break;
    default:
      break;
  }
}

int pikoracle_block_transform(int raw, int dir, const float* in,
                              float* out) {
  init_tables();
  if (raw < 0 || raw > 8) return -1;
  const int n = 8 * covered(raw);
  if (dir == 0) {
    to_pixels(raw, in, 64 * covered(raw), out, n);
  } else {
    from_pixels(raw, in, n, out, 64 * covered(raw));
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Reinterpreting transforms between DC-sized and coefficient-sized         */
/* grids, pik/ac_strategy.cc:272-315, and the AcStrategy methods using      */
/* them, pik/ac_strategy.cc:850-1064.                                       */

/* ReinterpretingDCT<dctn, n, n>. */
static void reinterp_dct(int dctn, int n, const float* in, size_t is,
                         float* out, size_t os) {
  float block[64], coef[64];
  for (int y = 0; y < n; ++y) {
    for (int x = 0; x < n; ++x) block[y * n + x] = in[y * is + x];
  }
  dct_2d(n, block, n, coef, n);
  for (int y = 0; y < n; ++y) {
    for (int x = 0; x < n; ++x) {
      out[y * os + x] =
          coef[y * n + x] * total_scale(n, x, y) * inv_total_scale(dctn, x, y);
    }
  }
}

/* ReinterpretingIDCT<dctn, lf, n>. */
static void reinterp_idct(int dctn, int lf, int n, const float* in, size_t is,
                          float* out, size_t os) {
  float block[64] = {0};
  for (int y = 0; y < lf; ++y) {
    for (int x = 0; x < lf; ++x) {
      block[y * n + x] =
          in[y * is + x] * inv_total_scale(n, x, y) * total_scale(dctn, x, y);
    }
  }
  idct_2d(n, block, n, out, os);
}

/* LowestFrequenciesFromDC, pik/ac_strategy.cc:850-887. */
static void llf_from_dc(int raw, const float* dc, size_t ds, float* llf,
                        size_t ls) {
  if (raw == kDCT16) {
    reinterp_dct(16, 2, dc, ds, llf, ls);
  } else if (raw == kDCT32) {
    reinterp_dct(32, 4, dc, ds, llf, ls);
  } else {
    llf[0] = dc[0];
  }
}

/* DC2x2FromLowestFrequencies, pik/ac_strategy.cc:928-969. */
static void dc2x2_from_llf(int raw, const float* llf, size_t ls, float* out,
                           size_t os) {
  if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
    reinterp_idct(8, 1, 2, llf, 0, out, os);
  } else if (raw == kDCT16) {
    reinterp_idct(16, 2, 4, llf, ls, out, os);
  } else if (raw == kDCT32) {
    reinterp_idct(32, 4, 8, llf, ls, out, os);
  } else {
    out[0] = llf[0];
    out[1] = llf[0];
    out[os] = llf[0];
    out[os + 1] = llf[0];
  }
}

/* DC2x2FromLowFrequencies, pik/ac_strategy.cc:971-1017; `block` points to
 * the transform's first storage row (coefficient F[y][x] at y * 8s + x). */
static void dc2x2_from_lf(int raw, const float* block, float* out, size_t os) {
  if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
    reinterp_idct(8, 2, 2, block, 8, out, os);
  } else if (raw == kDCT16) {
    reinterp_idct(16, 4, 4, block, 16, out, os);
  } else if (raw == kDCT32) {
    reinterp_idct(32, 8, 8, block, 32, out, os);
  } else {
    float d[4];
    hadamard_dc(block, d);
    out[0] = d[0];
    out[1] = d[1];
    out[os] = d[2];
    out[os + 1] = d[3];
  }
}

/* LowFrequenciesFromDC2x2, pik/ac_strategy.cc:1019-1064 (in place
 * allowed). */
static void lf_from_dc2x2(int raw, float* p, size_t ps) {
  if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
    reinterp_dct(8, 2, p, ps, p, ps);
  } else if (raw == kDCT16) {
    reinterp_dct(16, 4, p, ps, p, ps);
  } else if (raw == kDCT32) {
    reinterp_dct(32, 8, p, ps, p, ps);
  } else {
    const float b00 = p[0], b01 = p[1], b10 = p[ps], b11 = p[ps + 1];
    p[0] = (b00 + b01 + b10 + b11) * 0.25f;
    p[1] = (b00 + b01 - b10 - b11) * 0.25f;
    p[ps] = (b00 - b01 + b10 - b11) * 0.25f;
    p[ps + 1] = (b00 - b01 - b10 + b11) * 0.25f;
  }
}

/* ------------------------------------------------------------------------ */
/* 3x3 symmetric convolution with mirrored borders, in the operation order  */
/* of strategy::Symmetric3 (pik/convolve.h:586-625): w0 centre, w1 edges,   */
/* w2 corners.                                                              */

static void symmetric3(const float* in, int w, int h, const float wt[3],
                       float* out) {
  for (int y = 0; y < h; ++y) {
    const float* rt = in + mirror(y - 1, h) * (size_t)w;
    const float* rm = in + (size_t)y * w;
    const float* rb = in + mirror(y + 1, h) * (size_t)w;
    for (int x = 0; x < w; ++x) {
      const long xl = mirror(x - 1, w), xr = mirror(x + 1, w);
      const float sum_tb = rt[x] + rb[x];
      const float mul0 = rm[x] * wt[0];
      const float sum_lr = rm[xl] + rm[xr];
      const float x1 = sum_tb + sum_lr;
      const float mul1 = x1 * wt[1] + mul0;
      const float sum_t2 = rt[xl] + rt[xr];
      const float sum_b2 = rb[xl] + rb[xr];
      const float x2 = sum_t2 + sum_b2;
      out[(size_t)y * w + x] = x2 * wt[2] + mul1;
    }
  }
}

/* Gaborish weights, pik/gaborish.cc:16-85: unnormalised (1, wu1, wu2)
 * divided by wu0 + 4 (wu1 + wu2). */
static void gaborish_weights(int strength, float w[3]) {
  const float wu0 = 1.0f;
  float wu1, wu2;
  if (strength == 1) {
    wu1 = (float)0.056007960760453189;
    wu2 = (float)0.045899074552453879;
  } else if (strength == 2) {
    const float x = 0.75f;
    wu1 = (float)(x * 0.13959942428275746);
    wu2 = (float)(x * 0.074240717189152386);
  } else if (strength == 3) {
    const float x = 0.875f;
    wu1 = (float)(x * 0.11501538179658321);
    wu2 = (float)(x * 0.089979079587015454);
  } else {
    wu1 = (float)0.11501538179658321;
    wu2 = (float)0.089979079587015454;
  }
  const float mul = (float)(1.0 / (wu0 + 4 * (wu1 + wu2)));
  w[0] = wu0 * mul;
  w[1] = wu1 * mul;
  w[2] = wu2 * mul;
}

static void convolve3(const float* const in[3], int w, int h,
                      const float wt[3], float* const out[3]) {
  for (int c = 0; c < 3; ++c) symmetric3(in[c], w, h, wt, out[c]);
}

/* ------------------------------------------------------------------------ */
/* Edge-preserving filter, pik/epf_target.cc (scalar target).               */

/* WeightFast (pik/epf_target.cc:304-376): approximates exp(-s * sad) by
 * assembling a float from an integer exponent. */
static float weight_fast(int sad, int mul) {
  const int clamped = sad < 507 ? sad : 507;
  const int16_t prescaled = (int16_t)(clamped << 6);
  const int16_t unbiased = (int16_t)(((int32_t)prescaled * mul) >> 16);
  const int16_t biased = (int16_t)(unbiased + (127 << 7));
  const uint32_t bits = (uint32_t)(uint16_t)biased << 16;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

/* MulTable::Init (pik/epf_target.cc:379-421): for every fourth sigma the
 * first multiplier whose weight at SAD = sigma / 4 exceeds 0.5, linearly
 * interpolated in between (truncated to int). */
static void init_weight_table(void) {
  const int gap = 4;
  int mul = -32768;
  for (int sigma = kMinSigma; sigma <= kMaxSigma; sigma += gap) {
    for (; mul < 0; ++mul) {
      if (weight_fast(sigma >> 2, mul) > 0.5f) break;
    }
    g_mul_table[sigma] = mul;
  }
  for (int sigma = kMinSigma; sigma < kMaxSigma; sigma += gap) {
    const float step =
        (g_mul_table[sigma + gap] - g_mul_table[sigma]) / (float)gap;
    for (int i = 1; i < gap; ++i) {
      g_mul_table[sigma + i] = (int)(g_mul_table[sigma] + i * step);
    }
  }
}

/* SigmaFromQuant with the baseline LUTs, pik/epf_target.cc:1072-1111,
 * 1161-1196. */
static int sigma_from_quant(float signal, float stretch, int lut_id) {
  static const float lut[32] = {
      (float)1.9815775622811198, (float)1.9715084740908622,
      (float)1.6819963065873933, (float)1.2146133632942862,
      (float)1.0395364091521881, (float)0.93552327583169714,
      (float)0.68568655651684773, (float)0.51174440217871964,
      (float)0.36397262821018583, (float)0.31621830414136975,
      (float)0.30262954326557712, (float)0.246314237855494,
      (float)0.21617524864418683, (float)0.10, (float)0.05, (float)0.0,
      (float)(1.9815775622811198 / 2.0), (float)(1.9715084740908622 / 2.0),
      (float)(1.6819963065873933 / 2.0), (float)(1.2146133632942862 / 2.0),
      (float)(1.0395364091521881 / 2.0), (float)(0.93552327583169714 / 2.0),
      (float)(0.68568655651684773 / 2.0), (float)(0.51174440217871964 / 2.0),
      (float)(0.36397262821018583 / 2.0), (float)(0.31621830414136975 / 2.0),
      (float)(0.30262954326557712 / 2.0), (float)(0.246314237855494 / 2.0),
      (float)(0.21617524864418683 / 2.0), (float)(0.10 / 2.0),
      (float)(0.05 / 2.0), (float)(0.0 / 2.0)};
  const float* l = lut + 16 * lut_id;
  const float min_signal = 0.022156f;
  const float max_signal = (float)0.531738;
  const float mul_signal = (float)29.435892;
  float unscaled;
  if (signal <= min_signal) {
    unscaled = l[0];
  } else if (signal >= max_signal) {
    unscaled = l[15];
  } else {
    const float pos = (signal - min_signal) * mul_signal;
    const int64_t trunc = (int64_t)pos;
    const float frac = pos - (float)trunc;
    unscaled = frac * l[trunc + 1] + (1.0f - frac) * l[trunc];
  }
  const float bias = (float)0.5182760822018414;
  const int sigma = (int)(unscaled * stretch + bias);
  return sigma < kMaxSigma ? sigma : kMaxSigma;
}

/* Per-block inputs of the filter. */
typedef struct {
  const int32_t* qf; size_t qf_stride;  /* NULL: constant qf_const */
  int qf_const;
  const uint8_t* acs; size_t acs_stride;  /* NULL: all DCT8 */
  const uint8_t* lut; size_t lut_stride;  /* NULL: all 0 */
} EpfBlocks;

/* AdaptiveFilter (pik/epf_target.cc:1113-1277) on a w x h image (multiples
 * of 8): `in` is averaged, `guide_src` gives the 8-bit guide after the
 * global min/max stretch (pik/epf_target.cc:980-1070). */
static int epf(const float* const in[3], const float* const guide_src[3],
               int w, int h, const EpfBlocks* blk, float* const out[3]) {
  const int pw = w + 2 * kBorder, ph = h + 2 * kBorder;
  float* pin = (float*)malloc(sizeof(float) * 3 * pw * ph);
  uint8_t* gd = (uint8_t*)malloc(3 * (size_t)pw * ph);
  if (!pin || !gd) {
    free(pin);
    free(gd);
    return fail(PIKGPU_ENOMEM, "out of memory");
  }
  float mn[3], mx[3];
  for (int c = 0; c < 3; ++c) {
    mn[c] = mx[c] = guide_src[c][0];
    for (size_t i = 0; i < (size_t)w * h; ++i) {
      const float v = guide_src[c][i];
      if (v < mn[c]) mn[c] = v;
      if (mx[c] < v) mx[c] = v;
    }
  }
  float all_min = mn[0], all_max = mx[0];
  for (int c = 1; c < 3; ++c) {
    if (mn[c] < all_min) all_min = mn[c];
    if (all_max < mx[c]) all_max = mx[c];
  }
  const float stretch = all_min == all_max ? 1.f : 255.0f / (all_max - all_min);
  const float range = all_max - all_min;
  const float mul = range == 0.0f ? 1.0f : 255.0f / range;
  for (int c = 0; c < 3; ++c) {
    for (int y = 0; y < ph; ++y) {
      const long sy = mirror(y - kBorder, h);
      for (int x = 0; x < pw; ++x) {
        const long sx = mirror(x - kBorder, w);
        pin[((size_t)c * ph + y) * pw + x] = in[c][sy * w + sx];
        const float scaled = (guide_src[c][sy * w + sx] - all_min) * mul;
        gd[((size_t)c * ph + y) * pw + x] = (uint8_t)(int32_t)scaled;
      }
    }
  }
  const float quant_scale = 0.039324273f;
  for (int by = 0; by < h / 8; ++by) {
    for (int bx = 0; bx < w / 8; ++bx) {
      const float ac_q = (float)(blk->qf ? blk->qf[by * blk->qf_stride + bx]
                                         : blk->qf_const);
      const int raw = blk->acs ? blk->acs[by * blk->acs_stride + bx] >> 4 : 0;
      const int lut_id = blk->lut ? blk->lut[by * blk->lut_stride + bx] : 0;
      const float quant = ac_q * ar_quant_scale(raw) * quant_scale;
      const int sigma = sigma_from_quant(quant, stretch, lut_id);
      for (int iy = 0; iy < 8; ++iy) {
        const int y = by * 8 + iy;
        for (int ix = 0; ix < 8; ++ix) {
          const int x = bx * 8 + ix;
          if (sigma < kMinSigma) {
            for (int c = 0; c < 3; ++c) {
              out[c][(size_t)y * w + x] =
                  pin[((size_t)c * ph + y + kBorder) * pw + x + kBorder];
            }
            continue;
          }
          const int m = g_mul_table[sigma];
          int sad[56];
          for (int i = 0; i < 56; ++i) sad[i] = 0;
          for (int c = 0; c < 3; ++c) {
            const uint8_t* g = gd + (size_t)c * ph * pw;
            int i = 0;
            for (int cy = -3; cy <= 3; ++cy) {
              for (int cx = -3; cx <= 4; ++cx, ++i) {
                int s = 0;
                for (int py = -1; py <= 1; ++py) {
                  const uint8_t* rr = g + (size_t)(y + kBorder + py) * pw;
                  const uint8_t* rw =
                      g + (size_t)(y + kBorder + cy + py) * pw;
                  for (int px = -1; px <= 2; ++px) {
                    const int d = (int)rr[x + kBorder + px] -
                                  (int)rw[x + kBorder + cx + px];
                    s += d < 0 ? -d : d;
                  }
                }
                /* channels combined by their maximum
                 * (pik/epf_target.cc:625-660) */
                if (s > sad[i]) sad[i] = s;
              }
            }
          }
          float wts[56];
          for (int i = 0; i < 56; ++i) {
            const float wv = weight_fast(sad[i], m);
            wts[i] = wv >= 0.05f ? wv : 0.0f;
          }
          for (int c = 0; c < 3; ++c) {
            float ws = 0.0f, sw = 0.0f;
            int i = 0;
            for (int cy = -3; cy <= 3; ++cy) {
              const float* row =
                  pin + ((size_t)c * ph + y + kBorder + cy) * pw + kBorder;
              for (int cx = -3; cx <= 4; ++cx, ++i) {
                ws += row[x + cx] * wts[i];
                sw += wts[i];
              }
            }
            out[c][(size_t)y * w + x] = ws / sw;
          }
        }
      }
    }
  }
  free(pin);
  free(gd);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Frame state.                                                              */

typedef struct {
  const pikgpu_frame* f;
  int bx, by, px, py, tx;
  float inv_global_scale, inv_quant_dc;
  const float* table;
  const uint32_t* offsets;
  /* passes after the frame's own AC: extra[3 * p + c], rows of extra_stride */
  int num_extra;
  const int16_t* const* extra;
  size_t extra_stride;
} Ctx;

static const float* dequant_matrix(const Ctx* k, int table, int kind,
                                   int c) {
  return k->table + k->offsets[(table * 7 + kind) * 3 + c];
}

/* AdjustQuantBias, vector form (pik/quantizer.h:76-88). */
static float adjust_quant_bias(int c, float q) {
  static const float kOneBias[3] = {0.05465007330715401f, 0.07005449891748593f,
                                    0.049935103337343655f};
  const float aq = fabsf(q);
  if (aq >= 1.5f) return q - 0.145f / q;
  if (aq >= 0.5f) return q < 0 ? -(1 - kOneBias[c]) : 1 - kOneBias[c];
  return q;
}

/* Dequant<true>::DoAC (pik/compressed_image.cc:586-693) over the whole
 * frame: coefficient images ac[c] of by rows x bx*64 floats. */
static void dequant_ac(const Ctx* k, float* const ac[3]) {
  const pikgpu_frame* f = k->f;
  const size_t row = (size_t)k->bx * 64;
  for (int by = 0; by < k->by; ++by) {
    for (int bx = 0; bx < k->bx; ++bx) {
      const int qf = f->quant_field[by * f->quant_field_stride + bx];
      const float scaled = qf == 0 ? 1E10f : k->inv_global_scale / qf;
      const int tile_x = bx / 8, tile_y = by / 8;
      const int cf = f->dequant_cf[tile_y * f->dequant_cf_stride + tile_x];
      const int table = f->dequant_map[cf * 256 + qf - 1];
      const uint8_t a = f->ac_strategy[by * f->ac_strategy_stride + bx];
      const int kind = quant_kind(a >> 4), idx = a & 15;
      const float x_cc =
          1.0f * (f->ytox_map[tile_y * f->ytox_stride + tile_x] - 128) *
          (1.0f / 256);
      const float b_cc =
          1.0f * (f->ytob_map[tile_y * f->ytob_stride + tile_x] - 0) *
          (1.0f / 128);
      const float* m[3];
      for (int c = 0; c < 3; ++c) m[c] = dequant_matrix(k, table, kind, c) + idx * 64;
      for (int i = 0; i < 64; ++i) {
        const size_t at = (size_t)bx * 64 + i;
        /* Dequant<first>::DoAC per pass (pik/compressed_image.cc:575-703):
         * the first pass stores, later ones add their own dequantised
         * value to what is there. */
        for (int p = 0; p <= k->num_extra; ++p) {
          float q[3];
          for (int c = 0; c < 3; ++c) {
            q[c] = p == 0 ? (float)f->ac[c][by * f->ac_stride + at]
                          : (float)k->extra[3 * (p - 1) + c]
                                           [by * k->extra_stride + at];
          }
          const float xm = m[0][i] * scaled, ym = m[1][i] * scaled,
                      bm = m[2][i] * scaled;
          const float dx = adjust_quant_bias(0, q[0]) * xm;
          const float dy = adjust_quant_bias(1, q[1]) * ym;
          const float db = adjust_quant_bias(2, q[2]) * bm;
          const float v[3] = {x_cc * dy + dx, dy, b_cc * dy + db};
          for (int c = 0; c < 3; ++c) {
            float* dst = &ac[c][by * row + at];
            *dst = p == 0 ? v[c] : v[c] + *dst;
          }
        }
      }
    }
  }
}

/* DequantDC (pik/compressed_dc.cc:181-230) and AdaptiveDCReconstruction
 * (pik/adaptive_reconstruction.cc:300-332). */
static int reconstruct_dc(const Ctx* k, float* const dcq[3],
                          float* const dc[3]) {
  const pikgpu_frame* f = k->f;
  float mul_dc[3];
  for (int c = 0; c < 3; ++c) {
    mul_dc[c] = dequant_matrix(k, 0, 0, c)[0] * k->inv_quant_dc;
  }
  const float ytox_dc = 1.0f * (f->ytox_dc - 128) * (1.0f / 256);
  const float ytob_dc = 1.0f * (f->ytob_dc - 0) * (1.0f / 128);
  const size_t n = (size_t)k->bx * k->by;
  for (int y = 0; y < k->by; ++y) {
    for (int x = 0; x < k->bx; ++x) {
      const size_t i = (size_t)y * k->bx + x;
      const float yv = (float)f->dc[1][y * f->dc_stride + x] * mul_dc[1];
      dcq[1][i] = yv;
      dcq[0][i] = ytox_dc * yv + (float)f->dc[0][y * f->dc_stride + x] * mul_dc[0];
      dcq[2][i] = ytob_dc * yv + (float)f->dc[2][y * f->dc_stride + x] * mul_dc[2];
    }
  }
  /* PadImageToMultiple (pik/image.cc:45-75): replicate the last column and
   * row up to a multiple of 8. */
  const int pw = (int)div_ceil(k->bx, 8) * 8, ph = (int)div_ceil(k->by, 8) * 8;
  float* pad[3];
  float* den[3];
  float* mem = (float*)malloc(sizeof(float) * 6 * pw * ph);
  if (!mem) return fail(PIKGPU_ENOMEM, "out of memory");
  for (int c = 0; c < 3; ++c) {
    pad[c] = mem + (size_t)c * pw * ph;
    den[c] = mem + (size_t)(3 + c) * pw * ph;
    for (int y = 0; y < ph; ++y) {
      const int sy = y < k->by ? y : k->by - 1;
      for (int x = 0; x < pw; ++x) {
        const int sx = x < k->bx ? x : k->bx - 1;
        pad[c][(size_t)y * pw + x] = dcq[c][(size_t)sy * k->bx + sx];
      }
    }
  }
  EpfBlocks b = {NULL, 0, f->quant_dc / 2, NULL, 0, NULL, 0};
  const float* cin[3] = {pad[0], pad[1], pad[2]};
  int rc = epf(cin, cin, pw, ph, &b, den);
  if (rc) {
    free(mem);
    return rc;
  }
  for (int c = 0; c < 3; ++c) {
    const float half_step =
        k->inv_quant_dc * dequant_matrix(k, 0, 0, c)[0] * 0.5f;
    for (int y = 0; y < k->by; ++y) {
      for (int x = 0; x < k->bx; ++x) {
        const float v = dcq[c][(size_t)y * k->bx + x];
        const float d = den[c][(size_t)y * pw + x];
        const float hi = v + half_step, lo = v - half_step;
        /* std::max(lo, std::min(d, hi)) */
        const float mn = hi < d ? hi : d;
        dc[c][(size_t)y * k->bx + x] = lo < mn ? mn : lo;
      }
    }
  }
  (void)n;
  free(mem);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-group predictions and IDCT: InitializeDecCache                        */
/* (pik/compressed_dc.cc:530-556), ReconOpsinImage                          */
/* (pik/compressed_image.cc:936-1106) with ComputeLlf, PredictLf,           */
/* UpdateLfForDecoder and AddPredictions (pik/ac_predictions.cc:263-665).   */

/* SourceCoord, pik/compressed_dc.h:46-49. */
static size_t source_coord(size_t cand, size_t size) {
  return cand == 0 ? 0 : (cand == size + 1 ? size - 1 : cand - 1);
}

/* ComputeUb4Kernel (pik/ac_predictions.cc:234-254) with GaussianKernel
 * (pik/gauss_blur.h:18-32) of radius 4. */
static void ub4_kernel(float k[3][4]) {
  const float sigma = (float)2.0007879236394901;
  float kern[9];
  const float scaler = (float)(-1.0 / (2 * sigma * sigma));
  double sum = 0.0;
  for (int i = -4; i <= 4; ++i) {
    const float v = expf(scaler * i * i);
    kern[i + 4] = v;
    sum += v;
  }
  for (int i = 0; i < 9; ++i) kern[i] = (float)(kern[i] / sum);
  for (int j = 0; j < 3; ++j) {
    for (int q = 0; q < 4; ++q) k[j][q] = 0.0f;
  }
  for (int q = 0; q < 4; ++q) {
    const int s0 = 4 - q, s1 = 8 - q;
    for (int j = 0; j < s0; ++j) k[0][q] += kern[j];
    for (int j = s0; j < s1; ++j) k[1][q] += kern[j];
    for (int j = s1; j < 9; ++j) k[2][q] += kern[j];
  }
}

static int acs_raw_at(const Ctx* k, int bx, int by) {
  return k->f->ac_strategy[by * k->f->ac_strategy_stride + bx] >> 4;
}
static int acs_first_at(const Ctx* k, int bx, int by) {
  return (k->f->ac_strategy[by * k->f->ac_strategy_stride + bx] & 15) == 0;
}

/* GrayXyb (pik/compressed_image.cc:94-209): the X and B of a grayscale
 * frame's DC from its Y, by 16 line pieces fitted over 1024 gray levels of
 * LinearToXyb (pik/opsin_image.cc:34-47, with ApproxCubeRoot,
 * pik/approx_cube_root.h). */
enum { kGrayM = 16 };
static double g_gray_sub, g_gray_mul;
static double g_gray_xs[kGrayM], g_gray_xc[kGrayM], g_gray_bs[kGrayM],
    g_gray_bc[kGrayM];
static int g_gray_init;

static float approx_cbrt(float y) {
  int32_t ix;
  memcpy(&ix, &y, 4);
  ix = 0x2a50f200 + ix / 3;
  float x;
  memcpy(&x, &ix, 4);
  const float third = 1.0f / 3.0f;
  x = third * (2.0f * x + y / (x * x));
  x = third * (2.0f * x + y / (x * x));
  return x;
}

static void linear_to_xyb(float r, float g, float b, float* x, float* y,
                          float* z) {
  static const double m[9] = {0.29956550340058319, 0.63373087833825936,
                              0.077705617820981968, 0.22158691104574774,
                              0.68491388044116142, 0.096254234043612538,
                              0.20062661225219422, 0.070366199217588729,
                              0.5571760754215358};
  static const double bias[3] = {0.26786006338144885, 0.24494032763907073,
                                 0.14255999980363571};
  float mixed[3];
  for (int c = 0; c < 3; ++c) {
    const float m0 = (float)(m[3 * c] / 255.0f);
    const float m1 = (float)(m[3 * c + 1] / 255.0f);
    const float m2 = (float)(m[3 * c + 2] / 255.0f);
    const float bc = (float)(bias[c] / 255.0f);
    float v = m0 * r + m1 * g + m2 * b + bc;
    v = 0.0f < v ? v : 0.0f;
    mixed[c] = approx_cbrt(v);
  }
  *x = (1.0f * mixed[0] - 1.0f * mixed[1]) * 0.5f;
  *y = (1.0f * mixed[0] + 1.0f * mixed[1]) * 0.5f;
  *z = mixed[2];
}

static void gray_regression(const float* x, const float* y, size_t size,
                            double* a, double* b) {
  double mx = 0, my = 0, mx2 = 0, my2 = 0, mxy = 0;
  for (size_t i = 0; i < size; i++) {
    const double inv = 1.0 / (i + 1);
    const double dx = x[i] - mx;
    const double xn = dx * inv;
    mx += xn;
    mx2 += dx * xn * i;
    const double dy = y[i] - my;
    const double yn = dy * inv;
    my += yn;
    my2 += dy * yn * i;
    mxy += i * xn * yn - mxy * inv;
  }
  const double sx = sqrt(mx2 / (size - 1));
  const double sy = sqrt(my2 / (size - 1));
  const double sumxy = mxy * size + my * mx * size;
  const double r = (sumxy - size * mx * my) / ((size - 1.0) * sx * sy);
  *b = r * sy / sx;
  *a = my - *b * mx;
}

static void init_gray(void) {
  if (g_gray_init) return;
  enum { kN = 1024 };
  static float gx[kN], gy[kN], gz[kN];
  for (int i = 0; i < kN; i++) {
    const float gray = (float)(256.0f * i / kN);
    linear_to_xyb(gray, gray, gray, &gx[i], &gy[i], &gz[i]);
  }
  const float mn = gy[0], mxv = gy[kN - 1];
  int m = 0, border[kGrayM + 1];
  for (int i = 0; i < kN; i++) {
    if (gy[i] >= gy[0] + (mxv - mn) * m / kGrayM) {
      border[m] = i;
      m++;
    }
  }
  border[kGrayM] = kN;
  g_gray_sub = mn;
  g_gray_mul = 1.0 / (mxv - mn);
  for (int i = 0; i < kGrayM; i++) {
    const size_t n = (size_t)(border[i + 1] - border[i]);
    gray_regression(gy + border[i], gx + border[i], n, &g_gray_xc[i],
                    &g_gray_xs[i]);
    gray_regression(gy + border[i], gz + border[i], n, &g_gray_bc[i],
                    &g_gray_bs[i]);
  }
  g_gray_init = 1;
}

/* GrayXyb::YToXyb (pik/compressed_image.cc:99-104). */
static void y_to_xb(float y, float* x, float* b) {
  int i = (int)((y - g_gray_sub) * g_gray_mul * kGrayM);
  i = i < 0 ? 0 : (i > kGrayM - 1 ? kGrayM - 1 : i);
  *x = (float)(y * g_gray_xs[i] + g_gray_xc[i]);
  *b = (float)(y * g_gray_bs[i] + g_gray_bc[i]);
}

static int reconstruct_group(const Ctx* k, int gx, int gy, float* const dc[3],
                             float* const ac[3], float* const idct[3]) {
  const pikgpu_frame* f = k->f;
  const int x0 = gx * kGroupBlocks, y0 = gy * kGroupBlocks;
  const int gw = (k->bx - x0) < kGroupBlocks ? k->bx - x0 : kGroupBlocks;
  const int gh = (k->by - y0) < kGroupBlocks ? k->by - y0 : kGroupBlocks;
  const int cw = gw + 2, ch = gh + 2;       /* DC cache with border */
  const int w2 = 2 * cw, h2 = 2 * ch;       /* 2x images */
  const size_t row = (size_t)k->bx * 64;
  float* mem = (float*)calloc((size_t)3 * (2 * cw * ch + 3 * w2 * h2),
                              sizeof(float));
  if (!mem) return fail(PIKGPU_ENOMEM, "out of memory");
  float *cdc[3], *llf[3], *tmp2x2[3], *lf2x2[3], *pred2x2[3];
  for (int c = 0; c < 3; ++c) {
    float* base = mem + (size_t)c * (2 * cw * ch + 3 * w2 * h2);
    cdc[c] = base;
    llf[c] = base + cw * ch;
    tmp2x2[c] = base + 2 * cw * ch;
    lf2x2[c] = tmp2x2[c] + w2 * h2;
    pred2x2[c] = lf2x2[c] + w2 * h2;
  }
  /* InitializeDecCache: DC with a one-block replicated border. */
  for (int c = 0; c < 3; ++c) {
    for (int y = 0; y < ch; ++y) {
      const size_t sy = source_coord(y + y0, k->by);
      for (int x = 0; x < cw; ++x) {
        const size_t sx = source_coord(x + x0, k->bx);
        cdc[c][y * cw + x] = dc[c][sy * k->bx + sx];
      }
    }
  }
  /* kGrayscaleOpt: GrayXyb::RestoreXB on the group's DC
   * (pik/compressed_image.cc:953-956). */
  if (is_gray(f)) {
    init_gray();
    for (int i = 0; i < cw * ch; ++i) y_to_xb(cdc[1][i], &cdc[0][i], &cdc[2][i]);
  }
  /* Strategy at cache position (x, y): border positions are DCT8. */
#define CACHE_RAW(x, y)                                                  \
  (((x) == 0 || (y) == 0 || (x) == cw - 1 || (y) == ch - 1)             \
       ? kDCT                                                            \
       : acs_raw_at(k, x0 + (x) - 1, y0 + (y) - 1))
#define CACHE_FIRST(x, y)                                                \
  (((x) == 0 || (y) == 0 || (x) == cw - 1 || (y) == ch - 1)             \
       ? 1                                                               \
       : acs_first_at(k, x0 + (x) - 1, y0 + (y) - 1))

  /* ComputeLlf (pik/ac_predictions.cc:380-405). */
  for (int c = 0; c < 3; ++c) {
    for (int y = 0; y < ch; ++y) {
      for (int x = 0; x < cw; ++x) {
        if (!CACHE_FIRST(x, y)) continue;
        llf_from_dc(CACHE_RAW(x, y), cdc[c] + y * cw + x, cw,
                    llf[c] + y * cw + x, cw);
      }
    }
  }
  /* PredictLf (pik/ac_predictions.cc:410-470). */
  if (f->predict_lf) {
    static const float kLfBlur[3] = {(float)0.41459272584128337,
                                     (float)0.25489157325704559,
                                     (float)0.046449679523692139};
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < ch; ++y) {
        for (int x = 0; x < cw; ++x) {
          if (!CACHE_FIRST(x, y)) continue;
          dc2x2_from_llf(CACHE_RAW(x, y), llf[c] + y * cw + x, cw,
                         tmp2x2[c] + 2 * y * w2 + 2 * x, w2);
        }
      }
      symmetric3(tmp2x2[c], w2, h2, kLfBlur, lf2x2[c]);
      for (int y = 0; y < ch; ++y) {
        for (int x = 0; x < cw; ++x) {
          if (!CACHE_FIRST(x, y)) continue;
          lf_from_dc2x2(CACHE_RAW(x, y), lf2x2[c] + 2 * y * w2 + 2 * x, w2);
        }
      }
    }
  }
  /* Border of pred2x2 (pik/compressed_image.cc:976-1039); the rest of
   * pred2x2 starts at zero. */
  if (f->predict_hf) {
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < ch; ++y) {
        for (int x = 0; x < cw; ++x) {
          if (!(x == 0 || y == 0 || x == cw - 1 || y == ch - 1)) continue;
          float* o = pred2x2[c] + 2 * y * w2 + 2 * x;
          if (f->predict_lf) {
            float block[64] = {0};
            block[0] = llf[c][y * cw + x];
            block[1] = lf2x2[c][2 * y * w2 + 2 * x + 1];
            block[8] = lf2x2[c][(2 * y + 1) * w2 + 2 * x];
            block[9] = lf2x2[c][(2 * y + 1) * w2 + 2 * x + 1];
            dc2x2_from_lf(kDCT, block, o, w2);
          } else {
            dc2x2_from_llf(kDCT, llf[c] + y * cw + x, cw, o, w2);
          }
        }
      }
    }
  }
  /* UpdateLfForDecoder (pik/ac_predictions.cc:560-635) for every block. */
  for (int c = 0; c < 3; ++c) {
    for (int by = 0; by < gh; ++by) {
      for (int bx = 0; bx < gw; ++bx) {
        if (!acs_first_at(k, x0 + bx, y0 + by)) continue;
        const int raw = acs_raw_at(k, x0 + bx, y0 + by);
        const int s = covered(raw);
        float* a = ac[c] + (size_t)(y0 + by) * row + (size_t)(x0 + bx) * 64;
        for (int y = 0; y < s; ++y) {
          for (int x = 0; x < s; ++x) {
            a[s * 8 * y + x] = llf[c][(by + 1 + y) * cw + bx + 1 + x];
          }
        }
        if (f->predict_lf) {
          const float* lp = lf2x2[c] + 2 * (by + 1) * w2 + 2 * (bx + 1);
          for (int y = 0; y < 2 * s; ++y) {
            for (int x = y < s ? s : 0; x < 2 * s; ++x) {
              a[y * s * 8 + x] += lp[y * w2 + x];
            }
          }
        }
        if (f->predict_hf) {
          dc2x2_from_lf(raw, a,
                        pred2x2[c] + 2 * (by + 1) * w2 + 2 * (bx + 1), w2);
        }
      }
    }
  }
  /* AddPredictions / UpSample4x4BlurDCT<true>
   * (pik/ac_predictions.cc:263-374, 652-665). */
  if (f->predict_hf) {
    float kern[3][4];
    ub4_kernel(kern);
    const int xs = 2 * gw, ys = 2 * gh;
    float* blur_x = (float*)malloc(sizeof(float) * (size_t)xs * 4 * (ys + 2));
    if (!blur_x) {
      free(mem);
      return fail(PIKGPU_ENOMEM, "out of memory");
    }
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < ys + 2; ++y) {
        const float* r = pred2x2[c] + (y + 1) * w2;
        float* o = blur_x + (size_t)y * xs * 4;
        for (int x = 0; x < xs; ++x) {
          const float v0 = r[x + 1], v1 = r[x + 2], v2 = r[x + 3];
          for (int ix = 0; ix < 4; ++ix) {
            o[4 * x + ix] = v0 * kern[0][ix] + v1 * kern[1][ix] + v2 * kern[2][ix];
          }
        }
      }
      for (int by = 0; by < gh; ++by) {
        for (int bx = 0; bx < gw; ++bx) {
          if (!acs_first_at(k, x0 + bx, y0 + by)) continue;
          const int raw = acs_raw_at(k, x0 + bx, y0 + by);
          if (!predicts_hf(raw)) continue;
          const int s = covered(raw), n = 8 * s;
          float block[32 * 32], coeffs[32 * 32];
          for (int idy = 0; idy < s; ++idy) {
            const float* r0 = blur_x + (size_t)2 * (by + idy) * xs * 4;
            const float* r1 = r0 + xs * 4;
            const float* r2 = r1 + xs * 4;
            const float* r3 = r2 + xs * 4;
            for (int idx = 0; idx < s; ++idx) {
              for (int ix = 0; ix < 8; ++ix) {
                const int col = (bx + idx) * 8 + ix;
                const float v0 = r0[col], v1 = r1[col], v2 = r2[col],
                            v3 = r3[col];
                for (int iy = 0; iy < 4; ++iy) {
                  const float va =
                      v0 * kern[0][iy] + v1 * kern[1][iy] + v2 * kern[2][iy];
                  const float vb =
                      v1 * kern[0][iy] + v2 * kern[1][iy] + v3 * kern[2][iy];
                  block[(idy * 8 + iy) * n + idx * 8 + ix] = va;
                  block[(idy * 8 + 4 + iy) * n + idx * 8 + ix] = vb;
                }
              }
            }
          }
          from_pixels(raw, block, n, coeffs, 64 * s);
          /* AddBlockExceptLFAndLLFTo<true> (pik/ac_predictions.cc:38-70):
           * everything but the 2s x 2s low-frequency corner. */
          float* a = ac[c] + (size_t)(y0 + by) * row + (size_t)(x0 + bx) * 64;
          for (int i = 0; i < n * n; ++i) {
            const int fi = i / n, fj = i % n;
            if (fi < 2 * s && fj < 2 * s) continue;
            a[(size_t)(i / (64 * s)) * row + i % (64 * s)] += coeffs[i];
          }
        }
      }
    }
    free(blur_x);
  }
  /* InverseIntegralTransform (pik/compressed_image.cc:869-895). */
  for (int c = 0; c < 3; ++c) {
    for (int by = 0; by < gh; ++by) {
      for (int bx = 0; bx < gw; ++bx) {
        if (!acs_first_at(k, x0 + bx, y0 + by)) continue;
        const int raw = acs_raw_at(k, x0 + bx, y0 + by);
        to_pixels(raw, ac[c] + (size_t)(y0 + by) * row + (size_t)(x0 + bx) * 64,
                  row, idct[c] + (size_t)(y0 + by) * 8 * k->px + (x0 + bx) * 8,
                  k->px);
      }
    }
  }
#undef CACHE_RAW
#undef CACHE_FIRST
  free(mem);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* AdaptiveReconstruction's clamp, pik/adaptive_reconstruction.cc:73-244,   */
/* 334-471: per transform, the DCT of (filtered - original) is scaled by    */
/* the smallest ratio (over X, Y, B) that keeps every coefficient within    */
/* half a quantisation step, transformed back and added to the original.   */

static void ar_clamp(const Ctx* k, const float* const orig[3],
                     float* const filt[3]) {
  const pikgpu_frame* f = k->f;
  const size_t w = (size_t)k->px;
  float dc_mul[3];
  for (int c = 0; c < 3; ++c) {
    dc_mul[c] = k->inv_quant_dc * dequant_matrix(k, 0, 0, c)[0];
  }
  for (int by = 0; by < k->by; ++by) {
    for (int bx = 0; bx < k->bx; ++bx) {
      if (!acs_first_at(k, bx, by)) continue;
      const int raw = acs_raw_at(k, bx, by);
      const int s = covered(raw), n = 8 * s;
      const int qac = f->quant_field[by * f->quant_field_stride + bx];
      const float inv_quant_ac = k->inv_global_scale / qac;
      const int cf = f->dequant_cf[(by / 8) * f->dequant_cf_stride + bx / 8];
      const int table = f->dequant_map[cf * 256 + qac - 1];
      float blocks[3][32 * 32], ratio[32 * 32];
      for (int i = 0; i < n * n; ++i) ratio[i] = 1.0f;
      for (int c = 0; c < 3; ++c) {
        const float* dm = dequant_matrix(k, table, quant_kind(raw), c);
        float diff[32 * 32];
        const size_t at = (size_t)by * 8 * w + (size_t)bx * 8;
        for (int y = 0; y < n; ++y) {
          for (int x = 0; x < n; ++x) {
            diff[y * n + x] = filt[c][at + y * w + x] - orig[c][at + y * w + x];
          }
        }
        from_pixels(raw, diff, n, blocks[c], 64 * s);
        for (int i = 0; i < n * n; ++i) {
          const int fi = i / n, fj = i % n;
          float interval;
          if (fi < s && fj < s) {
            /* ARLowestFrequencyScales (pik/ac_strategy.h:171-205) */
            interval = inv_total_scale(n, fj, fi) * dc_mul[c];
            if (s == 1) interval = 1.0f * dc_mul[c];
          } else {
            interval = dm[i] * inv_quant_ac;
          }
          const float upper = 0.5f * interval;
          const float lower = -0.5f * interval;
          const float corr = blocks[c][i];
          const float mx = lower < corr ? corr : lower;
          const float clamped = mx < upper ? mx : upper;
          uint32_t bits;
          memcpy(&bits, &corr, 4);
          const float divisor = bits == 0 ? 1E-7f : corr;
          const float r = clamped / divisor;
          ratio[i] = r < ratio[i] ? r : ratio[i];
        }
      }
      static const int kOrder[3] = {2, 0, 1};  /* B, X, Y */
      for (int oc = 0; oc < 3; ++oc) {
        const int c = kOrder[oc];
        float pix[32 * 32];
        for (int i = 0; i < n * n; ++i) blocks[c][i] = blocks[c][i] * ratio[i];
        to_pixels(raw, blocks[c], 64 * s, pix, n);
        const size_t at = (size_t)by * 8 * w + (size_t)bx * 8;
        for (int y = 0; y < n; ++y) {
          for (int x = 0; x < n; ++x) {
            filt[c][at + y * w + x] = pix[y * n + x] + orig[c][at + y * w + x];
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* BlockDictionary::Apply<true> (pik/block_dictionary.cc:346-482).           */

static void add_dots(const Ctx* k, float* const img[3]) {
  const pikgpu_frame* f = k->f;
  const float kBlockInvScale = 1.0f / 3.5f;
  static const float kXybRadius[3] = {0.0301006f, 0.4512295f, 0.4716444f};
  const size_t w = (size_t)k->px;
  for (size_t i = 0; i < f->num_dots; ++i) {
    const pikgpu_dot* d = &f->dots[i];
    if (d->transform) continue;
    for (int c = 0; c < 3; ++c) {
      for (uint32_t iy = 0; iy < d->ysize; ++iy) {
        for (uint32_t ix = 0; ix < d->xsize; ++ix) {
          const float val = kBlockInvScale * kXybRadius[c] *
                            d->pixels[c][iy * d->xsize + ix];
          img[c][(d->y + iy) * w + d->x + ix] += val;
        }
      }
    }
  }
  for (size_t i = 0; i < f->num_dots; ++i) {
    const pikgpu_dot* d = &f->dots[i];
    if (!d->transform) continue;
    const size_t xs = d->xsize, ys = d->ysize;
    const size_t bw = 2 * xs + 5;
    float block[3][21 * 21];
    memset(block, 0, sizeof(block));
    for (int c = 0; c < 3; ++c) {
      for (int iy = 0; iy < (int)(2 * ys) + 1; ++iy) {
        for (int ix = 0; ix < (int)(2 * xs) + 1; ++ix) {
          float val = 0.0f;
          for (int dy = -1; dy < 1; ++dy) {
            for (int dx = -1; dx < 1; ++dx) {
              int sy = (iy + dy) / 2;
              if (sy < 0) sy = 0;
              if (sy >= (int)ys) sy = (int)ys - 1;
              int sx = (ix + dx) / 2;
              if (sx < 0) sx = 0;
              if (sx >= (int)xs) sx = (int)xs - 1;
              val += d->pixels[c][sy * xs + sx];
            }
          }
          block[c][(iy + 2) * bw + (ix + 2)] = val * 0.25f;
        }
      }
    }
    /* BlockPosition holds dx, dy, width as int64 and xsize, ysize as
     * size_t; the arithmetic below keeps those types' conversions. */
    const int64_t pdx = d->dx, pdy = d->dy, pwidth = d->width;
    const float by00 = (float)d->y;
    const float bx00 = (float)d->x;
    const float bx01 = bx00 + (float)pdx;
    const float by01 = by00 + (float)pdy;
    const float xnorm = (float)sqrt((double)(pdx * pdx + pdy * pdy));
    const float ynorm = (float)pwidth * 0.5f;
    const float invxnorm = 1.0f / xnorm;
    const float invynorm = 1.0f / ynorm;
    const float deltax = (float)pdy * invxnorm * ynorm;
    const float deltay = (float)(-pdx) * invxnorm * ynorm;
    const float bx11 = bx01 + deltax;
    const float by11 = by01 + deltay;
    const float bx10 = bx00 + deltax;
    const float by10 = by00 + deltay;
    const float inv_det = 1.0f / ((float)pdx * deltay - (float)pdy * deltax);
    const float it[4] = {(float)xs * deltay * inv_det,
                         -(float)xs * deltax * inv_det,
                         -(float)ys * (float)pdy * inv_det,
                         (float)((uint64_t)ys * (uint64_t)pdx) * inv_det};
#define FMIN(a, b) ((b) < (a) ? (b) : (a))
#define FMAX(a, b) ((a) < (b) ? (b) : (a))
    int64_t min_x = (int64_t)FMIN(FMIN(bx00, bx01), FMIN(bx10, bx11));
    int64_t min_y = (int64_t)FMIN(FMIN(by00, by01), FMIN(by10, by11));
    int64_t max_x = (int64_t)(FMAX(FMAX(bx00, bx01), FMAX(bx10, bx11)) + 1);
    int64_t max_y = (int64_t)(FMAX(FMAX(by00, by01), FMAX(by10, by11)) + 1);
#undef FMIN
#undef FMAX
    if (min_x < 0) min_x = 0;
    if (min_y < 0) min_y = 0;
    if (max_x > k->px) max_x = k->px;
    if (max_y > k->py) max_y = k->py;
    const float margin_x = 0.3f * invxnorm;
    const float margin_y = 0.3f * invynorm;
    for (int64_t iy = min_y; iy < max_y; ++iy) {
      for (int64_t ix = min_x; ix < max_x; ++ix) {
        const float x = (float)(size_t)ix - bx00 + 0.5f;
        const float y = (float)(size_t)iy - by00 + 0.5f;
        float ox = it[0] * x + it[1] * y;
        float oy = it[2] * x + it[3] * y;
        if (ox >= -margin_x && ox < xs + margin_x && oy >= -margin_y &&
            oy < ys + margin_y) {
          ox = 2 * ox + 2;
          oy = 2 * oy + 2;
          const int fx = (int)ox, cx = fx + 1;
          const float frx = ox - fx;
          const int fy = (int)oy, cy = fy + 1;
          const float fry = oy - fy;
          for (int c = 0; c < 3; ++c) {
            const float val =
                (frx * fry * block[c][cy * bw + cx] +
                 frx * (1.f - fry) * block[c][fy * bw + cx] +
                 (1.f - frx) * fry * block[c][cy * bw + fx] +
                 (1.f - frx) * (1.f - fry) * block[c][fy * bw + fx]) *
                kBlockInvScale * kXybRadius[c];
            img[c][iy * w + ix] += val;
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* OpsinToLinear / XybToRgb, pik/opsin_inverse.cc:39-123, with the inverse  */
/* of kOpsinAbsorbanceMatrix from Inv3x3Matrix (pik/linalg.h:153-170,       */
/* pik/opsin_params.{h,cc}).                                                 */

static void opsin_to_linear(float* const img[3], size_t n) {
  static const double kAbs[9] = {
      0.29956550340058319, 0.63373087833825936, 0.077705617820981968,
      0.22158691104574774, 0.68491388044116142, 0.096254234043612538,
      0.20062661225219422, 0.070366199217588729, 0.5571760754215358};
  static const double kBias[3] = {0.26786006338144885, 0.24494032763907073,
                                  0.14255999980363571};
  float m[9], t[9];
  for (int i = 0; i < 9; ++i) m[i] = (float)(kAbs[i] / 255.0f);
  t[0] = m[4] * m[8] - m[5] * m[7];
  t[1] = m[2] * m[7] - m[1] * m[8];
  t[2] = m[1] * m[5] - m[2] * m[4];
  t[3] = m[5] * m[6] - m[3] * m[8];
  t[4] = m[0] * m[8] - m[2] * m[6];
  t[5] = m[2] * m[3] - m[0] * m[5];
  t[6] = m[3] * m[7] - m[4] * m[6];
  t[7] = m[1] * m[6] - m[0] * m[7];
  t[8] = m[0] * m[4] - m[1] * m[3];
  const float idet = (float)(1.0 / (m[0] * t[0] + m[1] * t[3] + m[2] * t[6]));
  float inv[9];
  for (int i = 0; i < 9; ++i) inv[i] = t[i] * idet;
  float neg_bias[3];
  for (int c = 0; c < 3; ++c) neg_bias[c] = -(float)(kBias[c] / 255.0f);
  for (size_t i = 0; i < n; ++i) {
    const float ox = img[0][i], oy = img[1][i], ob = img[2][i];
    const float gr = 1.0f * (oy + ox);
    const float gg = 1.0f * (oy - ox);
    const float gb = ob;
    const float mr = gr * gr * gr + neg_bias[0];
    const float mg = gg * gg * gg + neg_bias[1];
    const float mb = gb * gb * gb + neg_bias[2];
    float lr = inv[0] * mr, lg = inv[3] * mr, lb = inv[6] * mr;
    const float tr = inv[1] * mg, tg = inv[4] * mg, tb = inv[7] * mg;
    lr = inv[2] * mb + lr;
    lg = inv[5] * mb + lg;
    lb = inv[8] * mb + lb;
    img[0][i] = lr + tr;
    img[1][i] = lg + tg;
    img[2][i] = lb + tb;
  }
}

/* kGrayscaleOpt frames after FinalizeFrameDecoding: every channel set to
 * the luma of linear RGB, evaluated in double (pik/pik_frame.cc:1155-1166). */
static void force_gray(float* const img[3], size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const float gray = (float)(img[0][i] * 0.299 + img[1][i] * 0.587 +
                               img[2][i] * 0.114);
    img[0][i] = img[1][i] = img[2][i] = gray;
  }
}

/* ------------------------------------------------------------------------ */

static void dump(float* const* stages, int s, float* const src[3], size_t w,
                 size_t h, size_t src_stride) {
  if (!stages) return;
  for (int c = 0; c < 3; ++c) {
    float* dst = stages[3 * s + c];
    if (!dst) continue;
    for (size_t y = 0; y < h; ++y) {
      memcpy(dst + y * w, src[c] + y * src_stride, w * sizeof(float));
    }
  }
}

static int check_frame(const pikgpu_frame* f) {
  if (!f) return fail(PIKGPU_EINVAL, "null frame");
  if (f->xsize == 0 || f->ysize == 0) return fail(PIKGPU_EINVAL, "empty image");
  if (f->global_scale < 1 || f->quant_dc < 1) {
    return fail(PIKGPU_EINVAL, "invalid quantizer scalars");
  }
  if (f->adaptive_reconstruction && !f->epf_enable_adaptive) {
    return fail(PIKGPU_EUNSUPPORTED, "fixed-sigma EPF");
  }
  if (f->flags & ~(uint32_t)PIKGPU_FLAG_GRAYSCALE_OPT) {
    return fail(PIKGPU_EUNSUPPORTED, "gradient map, noise or unknown flag");
  }
  const size_t bx = div_ceil(f->xsize, 8), by = div_ceil(f->ysize, 8);
  for (size_t y = 0; y < by; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      const int raw = f->ac_strategy[y * f->ac_strategy_stride + x] >> 4;
      if (raw > 8) return fail(PIKGPU_EINVAL, "invalid AC strategy");
      const int q = f->quant_field[y * f->quant_field_stride + x];
      if (q < 1 || q > 256) return fail(PIKGPU_EINVAL, "quant field range");
    }
  }
  return 0;
}

static int recon(const pikgpu_frame* f, int num_extra,
                 const int16_t* const* extra, size_t extra_stride,
                 float* const* stages, float* const out[3]);

int pikoracle_recon(const pikgpu_frame* f, float* const* stages,
                    float* const out[3]) {
  return recon(f, 0, NULL, 0, stages, out);
}

int pikoracle_recon_passes(const pikgpu_frame* f, int num_extra,
                           const int16_t* const* extra_ac,
                           size_t extra_stride, float* const* stages,
                           float* const out[3]) {
  if (num_extra < 0 || (num_extra > 0 && extra_ac == NULL)) {
    return fail(PIKGPU_EINVAL, "bad extra passes");
  }
  return recon(f, num_extra, extra_ac, extra_stride, stages, out);
}

static int recon(const pikgpu_frame* f, int num_extra,
                 const int16_t* const* extra, size_t extra_stride,
                 float* const* stages, float* const out[3]) {
  init_tables();
  int rc = check_frame(f);
  if (rc) return rc;
  Ctx k;
  k.f = f;
  k.num_extra = num_extra;
  k.extra = extra;
  k.extra_stride = extra_stride;
  k.bx = (int)div_ceil(f->xsize, 8);
  k.by = (int)div_ceil(f->ysize, 8);
  k.px = 8 * k.bx;
  k.py = 8 * k.by;
  k.tx = (int)div_ceil(f->xsize, 64);
  /* Quantizer::Decode (pik/quantizer.cc:63-71). */
  k.inv_global_scale = (float)(1.0 * 65536 / f->global_scale);
  k.inv_quant_dc = k.inv_global_scale / f->quant_dc;
  k.table = f->dequant_table;
  k.offsets = f->dequant_offsets;

  const size_t nb = (size_t)k.bx * k.by;
  const size_t np = (size_t)k.px * k.py;
  float* mem = (float*)calloc(6 * nb + 3 * nb * 64 + 9 * np, sizeof(float));
  if (!mem) return fail(PIKGPU_ENOMEM, "out of memory");
  float *dcq[3], *dc[3], *ac[3], *idct[3], *a[3], *b[3];
  for (int c = 0; c < 3; ++c) {
    dcq[c] = mem + c * nb;
    dc[c] = mem + (3 + c) * nb;
    ac[c] = mem + 6 * nb + c * nb * 64;
    idct[c] = mem + 6 * nb + 3 * nb * 64 + c * np;
    a[c] = mem + 6 * nb + 3 * nb * 64 + (3 + c) * np;
    b[c] = mem + 6 * nb + 3 * nb * 64 + (6 + c) * np;
  }
  if ((rc = reconstruct_dc(&k, dcq, dc))) goto done;
  dump(stages, PIKGPU_STAGE_DC_DEQUANT, dcq, k.bx, k.by, k.bx);
  dump(stages, PIKGPU_STAGE_DC, dc, k.bx, k.by, k.bx);

  dequant_ac(&k, ac);
  for (int gy = 0; gy < (int)div_ceil(k.by, kGroupBlocks); ++gy) {
    for (int gx = 0; gx < (int)div_ceil(k.bx, kGroupBlocks); ++gx) {
      if ((rc = reconstruct_group(&k, gx, gy, dc, ac, idct))) goto done;
    }
  }
  dump(stages, PIKGPU_STAGE_COEFFS, ac, (size_t)k.bx * 64, k.by,
       (size_t)k.bx * 64);
  dump(stages, PIKGPU_STAGE_IDCT, idct, k.px, k.py, k.px);

  /* FinalizeFrameDecoding (pik/compressed_image.cc:1110-1181). */
  float** cur = idct;
  if (f->adaptive_reconstruction) {
    const float* guide[3];
    if (f->gaborish) {
      float w[3];
      gaborish_weights(f->gaborish, w);
      convolve3((const float* const*)idct, k.px, k.py, w, a);
    }
    for (int c = 0; c < 3; ++c) {
      guide[c] = (f->gaborish && !f->epf_use_sharpened) ? a[c] : idct[c];
    }
    dump(stages, PIKGPU_STAGE_GUIDE, (float* const*)guide, k.px, k.py, k.px);
    EpfBlocks eb = {f->quant_field, f->quant_field_stride, 0,
                    f->ac_strategy, f->ac_strategy_stride,
                    f->sigma_lut_ids, f->sigma_lut_stride};
    if ((rc = epf((const float* const*)idct, guide, k.px, k.py, &eb, b))) {
      goto done;
    }
    dump(stages, PIKGPU_STAGE_EPF, b, k.px, k.py, k.px);
    ar_clamp(&k, (const float* const*)idct, b);
    cur = b;
  }
  dump(stages, PIKGPU_STAGE_AR, cur, k.px, k.py, k.px);
  add_dots(&k, cur);
  dump(stages, PIKGPU_STAGE_DOTS, cur, k.px, k.py, k.px);
  if (f->gaborish) {
    float w[3];
    float** dst = cur == b ? a : b;
    gaborish_weights(f->gaborish, w);
    convolve3((const float* const*)cur, k.px, k.py, w, dst);
    cur = dst;
  }
  dump(stages, PIKGPU_STAGE_GABORISH, cur, k.px, k.py, k.px);
  /* ShrinkTo + OpsinToLinear */
  for (int c = 0; c < 3; ++c) {
    for (uint32_t y = 0; y < f->ysize; ++y) {
      memcpy(out[c] + (size_t)y * f->xsize, cur[c] + (size_t)y * k.px,
             f->xsize * sizeof(float));
    }
  }
  opsin_to_linear(out, (size_t)f->xsize * f->ysize);
  if (is_gray(f)) force_gray(out, (size_t)f->xsize * f->ysize);
  dump(stages, PIKGPU_STAGE_LINEAR, out, f->xsize, f->ysize, f->xsize);
done:
  free(mem);
  return rc;
}

/* Downsampled decode (ReconOpsinImage / FinalizeFrameDecoding with
 * downsample = ds, pik/compressed_image.cc:869-934, 1144-1181): the
 * reconstruction up to the IDCT as at full size, each transform's pixels
 * averaged over ds x ds squares (the running sum in row order, times
 * 1 / ds^2; the squares are aligned to the transform, which is 8-aligned),
 * no AR and no Gaborish, dots refused (BlockDictionary::Apply checks there
 * are none), then ShrinkTo(ceil(xsize / ds), ceil(ysize / ds)) and
 * OpsinToLinear. `out`: 3 planes of ceil(xsize / ds) x ceil(ysize / ds). */
int pikoracle_recon_downsampled(const pikgpu_frame* f, int ds,
                                float* const out[3]) {
  init_tables();
  int rc = check_frame(f);
  if (rc) return rc;
  if (ds != 2 && ds != 4 && ds != 8) {
    return fail(PIKGPU_EINVAL, "downsampling factor must be 2, 4 or 8");
  }
  if (f->num_dots != 0) {
    return fail(PIKGPU_EUNSUPPORTED, "dots with downsampling");
  }
  Ctx k;
  k.f = f;
  k.num_extra = 0;
  k.extra = NULL;
  k.extra_stride = 0;
  k.bx = (int)div_ceil(f->xsize, 8);
  k.by = (int)div_ceil(f->ysize, 8);
  k.px = 8 * k.bx;
  k.py = 8 * k.by;
  k.tx = (int)div_ceil(f->xsize, 64);
  k.inv_global_scale = (float)(1.0 * 65536 / f->global_scale);
  k.inv_quant_dc = k.inv_global_scale / f->quant_dc;
  k.table = f->dequant_table;
  k.offsets = f->dequant_offsets;

  const size_t nb = (size_t)k.bx * k.by;
  const size_t np = (size_t)k.px * k.py;
  float* mem = (float*)calloc(6 * nb + 3 * nb * 64 + 3 * np, sizeof(float));
  if (!mem) return fail(PIKGPU_ENOMEM, "out of memory");
  float *dcq[3], *dc[3], *ac[3], *idct[3];
  for (int c = 0; c < 3; ++c) {
    dcq[c] = mem + c * nb;
    dc[c] = mem + (3 + c) * nb;
    ac[c] = mem + 6 * nb + c * nb * 64;
    idct[c] = mem + 6 * nb + 3 * nb * 64 + c * np;
  }
  if ((rc = reconstruct_dc(&k, dcq, dc))) goto done;
  dequant_ac(&k, ac);
  for (int gy = 0; gy < (int)div_ceil(k.by, kGroupBlocks); ++gy) {
    for (int gx = 0; gx < (int)div_ceil(k.bx, kGroupBlocks); ++gx) {
      if ((rc = reconstruct_group(&k, gx, gy, dc, ac, idct))) goto done;
    }
  }
  {
    const size_t ox = div_ceil(f->xsize, ds), oy = div_ceil(f->ysize, ds);
    const float mean_mul = 1.0f / (float)(ds * ds);
    for (int c = 0; c < 3; ++c) {
      for (size_t y = 0; y < oy; ++y) {
        for (size_t x = 0; x < ox; ++x) {
          float sum = 0.0f;
          for (int iy = 0; iy < ds; ++iy) {
            for (int ix = 0; ix < ds; ++ix) {
              sum += idct[c][(y * ds + iy) * k.px + x * ds + ix];
            }
          }
          out[c][y * ox + x] = sum * mean_mul;
        }
      }
    }
    opsin_to_linear(out, ox * oy);
    if (is_gray(f)) force_gray(out, ox * oy);
  }
done:
  free(mem);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Encoder-side quantization, as RoundtripImage runs it                      */
/* (pik/adaptive_quantization.cc:580-651): InitializeFrameEncCache           */
/* (pik/compressed_image.cc:210-304) and ComputeCoefficients (:351-455) for  */
/* frames without LF / HF prediction, dots or the grayscale flag.            */

/* Quantizer::QuantizeBlockAC (pik/quantizer.h:220-285) of one 8x8
 * sub-block (ix, iy) of an xs x ys transform: the 64 positions interleave
 * with the other sub-blocks', and the rounding error is carried forward
 * in double along each row and down from the previous row. `in` and `out`
 * index the transform's coefficients as pos = y * 8 xs + x. */
static void quantize_sub_block(const float* in, const float* qm, float qac,
                               float thres, int xs, int ys, int ix, int iy,
                               int16_t* out) {
  double prev[8] = {0};
  double err = 0;
  for (int k = 0; k < 64; ++k) {
    if ((k & 7) == 0) {
      err = prev[0];
    } else {
      err += prev[k & 7];
    }
    const int x = xs * (k % 8) + ix, y = ys * (k / 8) + iy;
    const int pos = y * 8 * xs + x;
    float val = in[pos] * (qm[pos] * qac);
    if (err > 0) {
      if (val > 0) {
        val += 0.5 * err;
      } else {
        val -= 0.5 * err;
      }
      err = 0;
    }
    if (fabsf(val) > 1) err = 0;
    double v = (fabsf(val) < thres) ? 0 : (double)roundf(val);
    if (fabs(v) < fabs((double)val)) {
      err += fabs(v - val);
    } else {
      err -= fabs(v - val);
    }
    if (k == 0) err = 0;
    if ((k & 7) == 7) {
      prev[k & 7] = 1.0 * err;
      err = 0;
    } else {
      err *= 0.5;
      prev[k & 7] = err;
    }
    if (v > 32767) v = 32767;
    if (v < -32767) v = -32767;
    out[pos] = (int16_t)v;
  }
}

/* AdjustQuantBiasVar (pik/quantizer.h:66-74) of an integer. */
static float adjust_quant_bias_int(int c, int q) {
  static const float kOneBias[3] = {0.05465007330715401f, 0.07005449891748593f,
                                    0.049935103337343655f};
  if (q == 0) return 0;
  if (q == 1) return 1 - kOneBias[c];
  if (q == -1) return kOneBias[c] - 1;
  return q - 0.145f / q;
}

/* The encoder's side of the LF / HF predictions for one group
 * (ComputeCoefficients' "enc predictions", pik/compressed_image.cc:373-388):
 * PredictLfForEncoder (pik/ac_predictions.cc:477-557) with
 * ComputeDecoderBlockAnd2x2DC (:136-199) and ComputeDecoderCoefficients
 * (:73-113), then ComputePredictionResiduals (:637-650). `coef` holds the
 * frame's float coefficients with the LLF already copied in; `cdc` / `llf`
 * the group's DC cache and its LLF (cw x ch, one-block border). On return
 * the LF positions of every transform hold the residual against the LF
 * prediction and the HF positions the residual against the HF prediction. */
static int encoder_predictions(const Ctx* k, int x0, int y0, int gw, int gh,
                               float* const cdc[3], float* const llf[3],
                               float* const coef[3]) {
  static const float kZeroBias[3] = {0.65f, 0.6f, 0.7f};
  static const float kLfBlur[3] = {(float)0.41459272584128337,
                                   (float)0.25489157325704559,
                                   (float)0.046449679523692139};
  const pikgpu_frame* f = k->f;
  const int cw = gw + 2, ch = gh + 2, w2 = 2 * cw, h2 = 2 * ch;
  const size_t row = (size_t)k->bx * 64;
  const float gsf = (float)(f->global_scale * (1.0 / 65536));
  float* mem = (float*)calloc((size_t)9 * w2 * h2, sizeof(float));
  if (!mem) return fail(PIKGPU_ENOMEM, "out of memory");
  float *tmp2x2[3], *lf2x2[3], *dc2x2[3];
  for (int c = 0; c < 3; ++c) {
    tmp2x2[c] = mem + (size_t)(3 * c) * w2 * h2;
    lf2x2[c] = tmp2x2[c] + (size_t)w2 * h2;
    dc2x2[c] = lf2x2[c] + (size_t)w2 * h2;
  }
#define ENC_BORDER(x, y) ((x) == 0 || (y) == 0 || (x) == cw - 1 || (y) == ch - 1)
  /* PredictLf, as the decoder forms it from the same LLF. */
  if (f->predict_lf) {
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < ch; ++y) {
        for (int x = 0; x < cw; ++x) {
          const int border = ENC_BORDER(x, y);
          if (!border && !acs_first_at(k, x0 + x - 1, y0 + y - 1)) continue;
          const int raw = border ? kDCT : acs_raw_at(k, x0 + x - 1, y0 + y - 1);
          dc2x2_from_llf(raw, llf[c] + y * cw + x, cw,
                         tmp2x2[c] + 2 * y * w2 + 2 * x, w2);
        }
      }
      symmetric3(tmp2x2[c], w2, h2, kLfBlur, lf2x2[c]);
      for (int y = 0; y < ch; ++y) {
        for (int x = 0; x < cw; ++x) {
          const int border = ENC_BORDER(x, y);
          if (!border && !acs_first_at(k, x0 + x - 1, y0 + y - 1)) continue;
          const int raw = border ? kDCT : acs_raw_at(k, x0 + x - 1, y0 + y - 1);
          lf_from_dc2x2(raw, lf2x2[c] + 2 * y * w2 + 2 * x, w2);
        }
      }
    }
  }
  /* Per block of the cache, Y then X then B: remove the LF prediction,
   * quantize the block as the encoder will (the LF positions' values
   * depend only on the low-frequency inputs) and dequantize its LF
   * positions, add the prediction back and form the block's 2x2 DC. */
  for (int yc = 0; yc < ch; ++yc) {
    for (int xc = 0; xc < cw; ++xc) {
      const int border = ENC_BORDER(xc, yc);
      float ydec[64 * 16];
      if (border) {
        for (int ci = 0; ci < 3; ++ci) {
          const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
          float dcf[64] = {0};
          dcf[0] = cdc[c][yc * cw + xc];
          if (f->predict_lf) {
            dcf[1] = lf2x2[c][2 * yc * w2 + 2 * xc + 1];
            dcf[8] = lf2x2[c][(2 * yc + 1) * w2 + 2 * xc];
            dcf[9] = lf2x2[c][(2 * yc + 1) * w2 + 2 * xc + 1];
          }
          if (f->predict_hf) {
            dc2x2_from_lf(kDCT, dcf, dc2x2[c] + 2 * yc * w2 + 2 * xc, w2);
          }
        }
        continue;
      }
      const int x = x0 + xc - 1, y = y0 + yc - 1;
      if (!acs_first_at(k, x, y)) continue;
      const int raw = acs_raw_at(k, x, y), s = covered(raw), n = 64 * s * s;
      const int kind = quant_kind(raw);
      const int quant = f->quant_field[y * f->quant_field_stride + x];
      const int tx = x / 8, ty = y / 8;
      const int cf = f->dequant_cf[ty * f->dequant_cf_stride + tx];
      if (cf * 256 + quant >= 16 * 256) {
        free(mem);
        return fail(PIKGPU_EUNSUPPORTED, "quant value past the dequant map");
      }
      const int table = f->dequant_map[cf * 256 + quant];
      const float qac = gsf * quant;
      const float inv_qac = k->inv_global_scale / quant;
      const float* lp = lf2x2[0];
      for (int ci = 0; ci < 3; ++ci) {
        const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
        const float fac =
            c == 0 ? 1.0f * (f->ytox_map[ty * f->ytox_stride + tx] - 128) *
                         (1.0f / 256)
                   : 1.0f * (f->ytob_map[ty * f->ytob_stride + tx] - 0) *
                         (1.0f / 128);
        const float* dm = dequant_matrix(k, table, kind, c);
        float inv_m[64 * 16], dcf[64 * 16];
        int16_t q[64 * 16];
        /* the transform's first storage row holds its 2s x 2s corner */
        float* blk = coef[c] + (size_t)y * row + (size_t)x * 64;
        lp = lf2x2[c] + 2 * yc * w2 + 2 * xc;
        if (f->predict_lf) {
          for (int i = 0; i < 2 * s; ++i) {
            for (int j = i < s ? s : 0; j < 2 * s; ++j) {
              blk[i * 8 * s + j] -= lp[i * w2 + j];
            }
          }
        }
        for (int r = 0; r < s; ++r) {
          memcpy(dcf + r * 64 * s, coef[c] + (size_t)(y + r) * row +
                                       (size_t)x * 64,
                 sizeof(float) * 64 * s);
        }
        if (c != 1) {
          for (int i = 0; i < n; ++i) dcf[i] -= ydec[i] * fac;
        }
        for (int i = 0; i < n; ++i) inv_m[i] = 1.0f / dm[i];
        for (int iy = 0; iy < s; ++iy) {
          for (int ix = 0; ix < s; ++ix) {
            quantize_sub_block(dcf, inv_m, qac, kZeroBias[c], s, s, ix, iy, q);
          }
        }
        /* QuantizeRoundtripBlockCoefficients (pik/quantizer.h:317-348) of
         * positions 1, 8, 9 of every interleaved 8x8 sub-block */
        for (int iy = 0; iy < s; ++iy) {
          for (int ix = 0; ix < s; ++ix) {
            static const int kLf[3] = {1, 8, 9};
            for (int m = 0; m < 3; ++m) {
              const int kx = kLf[m] % 8, ky = kLf[m] / 8;
              const int pos = (s * ky + iy) * 8 * s + s * kx + ix;
              dcf[pos] = adjust_quant_bias_int(c, q[pos]) * dm[pos] * inv_qac;
            }
          }
        }
        if (c != 1) {
          for (int i = 0; i < n; ++i) dcf[i] += ydec[i] * fac;
        } else {
          memcpy(ydec, dcf, sizeof(float) * n);
        }
        if (f->predict_lf) {
          for (int i = 0; i < 2 * s; ++i) {
            const int start = i < s ? s : 0;
            for (int j = 0; j < start; ++j) dcf[i * 8 * s + j] = blk[i * 8 * s + j];
            for (int j = start; j < 2 * s; ++j) dcf[i * 8 * s + j] += lp[i * w2 + j];
          }
        }
        if (f->predict_hf) {
          dc2x2_from_lf(raw, dcf, dc2x2[c] + 2 * yc * w2 + 2 * xc, w2);
        }
      }
    }
  }
#undef ENC_BORDER
  /* ComputePredictionResiduals: UpSample4x4BlurDCT<false>
   * (pik/ac_predictions.cc:263-374): the same prediction the decoder adds,
   * subtracted outside each transform's 2s x 2s corner. */
  if (f->predict_hf) {
    float kern[3][4];
    ub4_kernel(kern);
    const int xs = 2 * gw, ys = 2 * gh;
    float* blur_x = (float*)malloc(sizeof(float) * (size_t)xs * 4 * (ys + 2));
    if (!blur_x) {
      free(mem);
      return fail(PIKGPU_ENOMEM, "out of memory");
    }
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < ys + 2; ++y) {
        const float* r = dc2x2[c] + (y + 1) * w2;
        float* o = blur_x + (size_t)y * xs * 4;
        for (int x = 0; x < xs; ++x) {
          const float v0 = r[x + 1], v1 = r[x + 2], v2 = r[x + 3];
          for (int ix = 0; ix < 4; ++ix) {
            o[4 * x + ix] = v0 * kern[0][ix] + v1 * kern[1][ix] + v2 * kern[2][ix];
          }
        }
      }
      for (int by = 0; by < gh; ++by) {
        for (int bx = 0; bx < gw; ++bx) {
          if (!acs_first_at(k, x0 + bx, y0 + by)) continue;
          const int raw = acs_raw_at(k, x0 + bx, y0 + by);
          if (!predicts_hf(raw)) continue;
          const int s = covered(raw), n = 8 * s;
          float block[32 * 32], coeffs[32 * 32];
          for (int idy = 0; idy < s; ++idy) {
            const float* r0 = blur_x + (size_t)2 * (by + idy) * xs * 4;
            const float* r1 = r0 + xs * 4;
            const float* r2 = r1 + xs * 4;
            const float* r3 = r2 + xs * 4;
            for (int idx = 0; idx < s; ++idx) {
              for (int ix = 0; ix < 8; ++ix) {
                const int col = (bx + idx) * 8 + ix;
                const float v0 = r0[col], v1 = r1[col], v2 = r2[col],
                            v3 = r3[col];
                for (int iy = 0; iy < 4; ++iy) {
                  const float va =
                      v0 * kern[0][iy] + v1 * kern[1][iy] + v2 * kern[2][iy];
                  const float vb =
                      v1 * kern[0][iy] + v2 * kern[1][iy] + v3 * kern[2][iy];
                  block[(idy * 8 + iy) * n + idx * 8 + ix] = va;
                  block[(idy * 8 + 4 + iy) * n + idx * 8 + ix] = vb;
                }
              }
            }
          }
          from_pixels(raw, block, n, coeffs, 64 * s);
          float* a = coef[c] + (size_t)(y0 + by) * row + (size_t)(x0 + bx) * 64;
          for (int i = 0; i < n * n; ++i) {
            const int fi = i / n, fj = i % n;
            if (fi < 2 * s && fj < 2 * s) continue;
            a[(size_t)(i / (64 * s)) * row + i % (64 * s)] -= coeffs[i];
          }
        }
      }
    }
    free(blur_x);
  }
  free(mem);
  return 0;
}

int pikoracle_quantize(const pikgpu_frame* f, const float* const* opsin,
                       int16_t* const* ac_out, int16_t* const* dc_out) {
  init_tables();
  if (!f || !opsin) return fail(PIKGPU_EINVAL, "null argument");
  if (f->num_dots || is_gray(f)) {
    return fail(PIKGPU_EUNSUPPORTED,
                "quantization of frames with dots or grayscale");
  }
  pikgpu_frame g = *f;
  const int bx = (int)div_ceil(f->xsize, 8), by = (int)div_ceil(f->ysize, 8);
  const size_t nb = (size_t)bx * by, row = (size_t)bx * 64;
  const size_t px = (size_t)bx * 8;
  int16_t* q16 = (int16_t*)calloc(3 * nb * 65, sizeof(int16_t));
  float* mem = (float*)calloc(3 * nb * 64 + 9 * nb, sizeof(float));
  int rc = 0;
  if (!q16 || !mem) {
    rc = fail(PIKGPU_ENOMEM, "out of memory");
    goto done;
  }
  {
    float *coef[3], *dcf[3], *dcq[3], *dc[3];
    int16_t *ac16[3], *dc16[3];
    for (int c = 0; c < 3; ++c) {
      coef[c] = mem + c * nb * 64;
      dcf[c] = mem + 3 * nb * 64 + c * nb;
      dcq[c] = mem + 3 * nb * 64 + (3 + c) * nb;
      dc[c] = mem + 3 * nb * 64 + (6 + c) * nb;
      ac16[c] = q16 + c * nb * 64;
      dc16[c] = q16 + 3 * nb * 64 + c * nb;
      g.ac[c] = ac16[c];
      g.dc[c] = dc16[c];
    }
    g.ac_stride = row;
    g.dc_stride = (size_t)bx;
    if ((rc = check_frame(&g))) goto done;
    Ctx k;
    memset(&k, 0, sizeof(k));
    k.f = &g;
    k.bx = bx;
    k.by = by;
    k.px = 8 * bx;
    k.py = 8 * by;
    k.tx = (int)div_ceil(f->xsize, 64);
    k.inv_global_scale = (float)(1.0 * 65536 / f->global_scale);
    k.inv_quant_dc = k.inv_global_scale / f->quant_dc;
    k.table = f->dequant_table;
    k.offsets = f->dequant_offsets;
    /* TransformFromPixels + DCFromLowestFrequencies
     * (pik/compressed_image.cc:234-247, pik/ac_strategy.cc:889-927). */
    for (int c = 0; c < 3; ++c) {
      for (int y = 0; y < by; ++y) {
        for (int x = 0; x < bx; ++x) {
          if (!acs_first_at(&k, x, y)) continue;
          const int raw = acs_raw_at(&k, x, y);
          float* cb = coef[c] + y * row + (size_t)x * 64;
          from_pixels(raw, opsin[c] + (size_t)y * 8 * px + (size_t)x * 8, px,
                      cb, row);
          float* d = dcf[c] + (size_t)y * bx + x;
          if (raw == kDCT16) {
            reinterp_idct(16, 2, 2, cb, 16, d, (size_t)bx);
          } else if (raw == kDCT32) {
            reinterp_idct(32, 4, 4, cb, 32, d, (size_t)bx);
          } else {
            d[0] = cb[0];
          }
        }
      }
    }
    /* QuantizeRoundtripDC of Y, ApplyColorCorrelationDC<false>,
     * QuantizeCoeffsDC (pik/compressed_image.cc:262-271,
     * pik/quantizer.cc:86-117, pik/color_correlation.cc:180-205). */
    {
      const float gsf = (float)(f->global_scale * (1.0 / 65536));
      float inv0[3], mul[3];
      for (int c = 0; c < 3; ++c) {
        const float m0 = dequant_matrix(&k, 0, 0, c)[0];
        inv0[c] = 1.0f / m0;
        mul[c] = m0 * k.inv_quant_dc;
      }
      const float qdc = gsf * f->quant_dc;
      const float ytox = 1.0f * (f->ytox_dc - 128) * (1.0f / 256);
      const float ytob = 1.0f * (f->ytob_dc - 0) * (1.0f / 128);
      for (size_t i = 0; i < nb; ++i) {
        const int qy = (int)roundf(dcf[1][i] * (inv0[1] * qdc));
        const float ydec = qy * mul[1];
        const float xv = dcf[0][i] - ytox * ydec;
        const float bv = dcf[2][i] - ytob * ydec;
        dc16[1][i] = (int16_t)qy;
        dc16[0][i] = (int16_t)(int)roundf(xv * (inv0[0] * qdc));
        dc16[2][i] = (int16_t)(int)roundf(bv * (inv0[2] * qdc));
      }
    }
    /* dc_dec: the decoder's DequantDC and AdaptiveDCReconstruction do the
     * same arithmetic as the encoder's (pik/compressed_image.cc:272-298). */
    if ((rc = reconstruct_dc(&k, dcq, dc))) goto done;
    /* Per group: InitializeEncCache's DC with a replicated border, then
     * ComputeLlf and CopyLlf (pik/ac_predictions.cc:380-405, 581-607). */
    for (int gy = 0; gy < (int)div_ceil(by, kGroupBlocks); ++gy) {
      for (int gx = 0; gx < (int)div_ceil(bx, kGroupBlocks); ++gx) {
        const int x0 = gx * kGroupBlocks, y0 = gy * kGroupBlocks;
        const int gw = bx - x0 < kGroupBlocks ? bx - x0 : kGroupBlocks;
        const int gh = by - y0 < kGroupBlocks ? by - y0 : kGroupBlocks;
        const int cw = gw + 2, ch = gh + 2;
        float* cmem = (float*)calloc((size_t)6 * cw * ch, sizeof(float));
        if (!cmem) {
          rc = fail(PIKGPU_ENOMEM, "out of memory");
          goto done;
        }
        float *cdc[3], *cllf[3];
        for (int c = 0; c < 3; ++c) {
          cdc[c] = cmem + (size_t)c * cw * ch;
          cllf[c] = cmem + (size_t)(3 + c) * cw * ch;
          for (int y = 0; y < ch; ++y) {
            const size_t sy = source_coord(y + y0, by);
            for (int x = 0; x < cw; ++x) {
              cdc[c][y * cw + x] = dc[c][sy * bx + source_coord(x + x0, bx)];
            }
          }
          /* ComputeLlf over the cache (border positions are DCT8) */
          for (int y = 0; y < ch; ++y) {
            for (int x = 0; x < cw; ++x) {
              const int border = x == 0 || y == 0 || x == cw - 1 || y == ch - 1;
              if (!border && !acs_first_at(&k, x0 + x - 1, y0 + y - 1)) continue;
              const int raw =
                  border ? kDCT : acs_raw_at(&k, x0 + x - 1, y0 + y - 1);
              llf_from_dc(raw, cdc[c] + y * cw + x, cw, cllf[c] + y * cw + x,
                          cw);
            }
          }
          for (int y = 0; y < gh; ++y) {
            for (int x = 0; x < gw; ++x) {
              if (!acs_first_at(&k, x0 + x, y0 + y)) continue;
              const int raw = acs_raw_at(&k, x0 + x, y0 + y), s = covered(raw);
              float* cb = coef[c] + (y0 + y) * row + (size_t)(x0 + x) * 64;
              for (int iy = 0; iy < s; ++iy) {
                for (int ix = 0; ix < s; ++ix) {
                  cb[s * 8 * iy + ix] = cllf[c][(y + 1 + iy) * cw + x + 1 + ix];
                }
              }
            }
          }
        }
        if (f->predict_lf || f->predict_hf) {
          rc = encoder_predictions(&k, x0, y0, gw, gh, cdc, cllf, coef);
        }
        free(cmem);
        if (rc) goto done;
      }
    }
    /* QuantizeRoundtripBlockAC of Y, UnapplyColorCorrelationAC, then
     * QuantizeBlockAC of every channel (pik/compressed_image.cc:396-453). */
    {
      static const float kZeroBias[3] = {0.65f, 0.6f, 0.7f};
      const float gsf = (float)(f->global_scale * (1.0 / 65536));
      float inv_m[64 * 16];
      int16_t qy[64 * 16];
      float ydec[64 * 16], vals[64 * 16];
      for (int y = 0; y < by; ++y) {
        for (int x = 0; x < bx; ++x) {
          if (!acs_first_at(&k, x, y)) continue;
          const int raw = acs_raw_at(&k, x, y), s = covered(raw);
          const int kind = quant_kind(raw);
          const int quant = f->quant_field[y * f->quant_field_stride + x];
          const int tx = x / 8, ty = y / 8;
          const int cf = f->dequant_cf[ty * f->dequant_cf_stride + tx];
          /* indexed by the quant value itself, as the encoder does */
          if (cf * 256 + quant >= 16 * 256) {
            rc = fail(PIKGPU_EUNSUPPORTED, "quant value past the dequant map");
            goto done;
          }
          const int table = f->dequant_map[cf * 256 + quant];
          const float qac = gsf * quant;
          const float inv_qac = k.inv_global_scale / quant;
          const int n = 64 * s * s;
          /* Y first: X and B subtract its dequantised values */
          for (int ci = 0; ci < 3; ++ci) {
            const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
            const float* dm = dequant_matrix(&k, table, kind, c);
            for (int i = 0; i < n; ++i) inv_m[i] = 1.0f / dm[i];
            /* the transform's coefficients as pos = y * 8s + x, with the
             * colour correlation of Y removed from X and B */
            for (int r = 0; r < s; ++r) {
              for (int i = 0; i < 64 * s; ++i) {
                const size_t at = (size_t)(y + r) * row + (size_t)x * 64 + i;
                const int pos = r * 64 * s + i;
                float v = coef[c][at];
                if (c != 1) {
                  const int ttx = (x + i / 64) / 8, tty = (y + r) / 8;
                  const float fac =
                      c == 0 ? 1.0f * (f->ytox_map[tty * f->ytox_stride + ttx] -
                                       128) * (1.0f / 256)
                             : 1.0f * (f->ytob_map[tty * f->ytob_stride + ttx] -
                                       0) * (1.0f / 128);
                  v = v - fac * ydec[pos];
                }
                vals[pos] = v;
              }
            }
            int16_t qbuf[64 * 16];
            int16_t* qout = c == 1 ? qy : qbuf;
            for (int iy = 0; iy < s; ++iy) {
              for (int ix = 0; ix < s; ++ix) {
                quantize_sub_block(vals, inv_m, qac, kZeroBias[c], s, s, ix,
                                   iy, qout);
              }
            }
            if (c == 1) {
              for (int i = 0; i < n; ++i) {
                ydec[i] = adjust_quant_bias_int(1, qy[i]) * dm[i] * inv_qac;
              }
            }
            for (int r = 0; r < s; ++r) {
              memcpy(ac16[c] + (size_t)(y + r) * row + (size_t)x * 64,
                     qout + r * 64 * s, sizeof(int16_t) * 64 * s);
            }
          }
        }
      }
    }
    for (int c = 0; c < 3; ++c) {
      if (ac_out && ac_out[c]) memcpy(ac_out[c], ac16[c], sizeof(int16_t) * nb * 64);
      if (dc_out && dc_out[c]) memcpy(dc_out[c], dc16[c], sizeof(int16_t) * nb);
    }
  }
done:
  free(q16);
  free(mem);
  return rc;
}
