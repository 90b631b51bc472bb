// Host-side constant tables of the reconstruction, computed once per context.
//
// Each routine restates how the reference derives the same numbers so that
// the device kernels see bit-identical constants:
//  - EPF weight multipliers: MulTable::Init (pik/epf_target.cc:379-421) with
//    the WeightFast bit trick (pik/epf_target.cc:304-376);
//  - Gaborish 3x3 weights (pik/gaborish.cc:9-90);
//  - LF prediction blur weights (pik/ac_predictions.cc:121-132);
//  - the 4x upsampling blur of the HF prediction (ComputeUb4Kernel,
//    pik/ac_predictions.cc:234-254, GaussianKernel, pik/gauss_blur.h:18-32);
//  - the scale factors of the reinterpreting DCTs (DCTTotalScale /
//    DCTInvTotalScale, pik/ac_strategy.cc:258-270) and of AR's lowest
//    frequencies (pik/ac_strategy.h:33-38);
#include <math.h>

#include <cmath>
#include <string.h>

#include <vector>

#include "common.h"

namespace pikgpu {

namespace {

// First eight entries of the DCT normalisation tables (pik/dct.h:44-130,
// 142-245), N = 2, 4, 8, 16, 32. These define the reference's coefficient
// scaling and must be the same numbers for the reinterpreted LF coefficients
// to match. Only indices < 8 are ever used (LF regions are at most 8x8).
const float kDctScales[5][8] = {
    {0.707106781186547524f, 0.707106781186547524f},
    {0.5f, 0.653281482438188264f, 0.5f, 0.270598050073098492f},
    {0.353553390593273762f, 0.254897789552079584f, 0.270598050073098492f,
     0.30067244346752264f, 0.353553390593273762f, 0.449988111568207852f,
     0.653281482438188264f, 1.28145772387075309f},
    {0.25f, 0.351850934381595615f, 0.346759961330536865f,
     0.33832950029358817f, 0.326640741219094132f, 0.311806253246667808f,
     0.293968900604839679f, 0.273300466750439372f},
    {0.176776695296636881f, 0.249698864051293098f, 0.248796181668049222f,
     0.247294127491195243f, 0.245196320100807612f, 0.242507813298635998f,
     0.239235083933052216f, 0.235386016295755195f}};
const float kIdctScales[5][8] = {
    {0.707106781186547524f, 0.707106781186547524f},
    {0.5f, 0.382683432365089772f, 0.5f, 0.923879532511286756f},
    {0.353553390593273762f, 0.490392640201615225f, 0.461939766255643378f,
     0.415734806151272619f, 0.353553390593273762f, 0.277785116509801112f,
     0.191341716182544886f, 0.0975451610080641339f},
    {0.25f, 0.177632042131274808f, 0.180239955501736978f,
     0.184731156892216368f, 0.191341716182544886f, 0.200444985785954314f,
     0.212607523691814112f, 0.228686034616512494f},
    {0.176776695296636881f, 0.125150749558799075f, 0.125604821547038926f,
     0.126367739974385915f, 0.127448894776039792f, 0.128861827480656137f,
     0.13062465373492222f, 0.132760647772446044f}};

int Log2Index(int n) {  // 2->0, 4->1, 8->2, 16->3, 32->4
  int i = 0;
  while ((2 << i) < n) ++i;
  return i;
}

int TrailingZeros(int k) {
  int v = 0;
  while ((k & 1) == 0) {
    k >>= 1;
    ++v;
  }
  return v;
}

// Relative L1 norm of a single-coefficient IDCT (pik/dct.h:140-245): the
// product cos(pi/8) cos(pi/16) ... with log2(N) - 1 - v2(k) factors.
double L1Norm(int n, int k) {
  if (k == 0) return 1.0;
  const int factors = Log2Index(n) - TrailingZeros(k);
  double p = 1.0;
  for (int i = 0; i < factors; ++i) p *= cos(M_PI / double(8 << i));
  return p;
}

// The reference stores the L1 norms and their inverses as float literals of
// the exact values; rounding the double value reproduces them, except the
// 32-point inverses whose literals carry a little extra error in the last
// digits, which still round to the same floats.
float L1NormF(int n, int k) { return static_cast<float>(L1Norm(n, k)); }
float L1NormInvF(int n, int k) { return static_cast<float>(1.0 / L1Norm(n, k)); }

// DCTTotalScale<N>(x, y) and DCTInvTotalScale<N>(x, y), evaluated in float
// left to right like the reference's constexpr functions.
float TotalScale(int n, int x, int y) {
  const int i = Log2Index(n);
  float v = static_cast<float>(n);
  v = v * kDctScales[i][x];
  v = v * kDctScales[i][y];
  v = v * L1NormInvF(n, x);
  v = v * L1NormInvF(n, y);
  return v;
}
float InvTotalScale(int n, int x, int y) {
  const int i = Log2Index(n);
  float v = static_cast<float>(n);
  v = v * kIdctScales[i][x];
  v = v * kIdctScales[i][y];
  v = v * L1NormF(n, x);
  v = v * L1NormF(n, y);
  return v;
}

float WeightFromBits(int sad, int mul) {
  const int clamped = sad < kClampedSAD ? sad : kClampedSAD;
  const int16_t prescaled = static_cast<int16_t>(clamped << 6);
  const int16_t unbiased =
      static_cast<int16_t>((int32_t(prescaled) * int32_t(mul)) >> 16);
  const int16_t biased = static_cast<int16_t>(unbiased + (127 << 7));
  const uint32_t bits = static_cast<uint32_t>(static_cast<uint16_t>(biased))
                        << 16;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

}  // namespace

void ComputeHostTables(HostTables* t) {
  memset(t->mul_table, 0, sizeof(t->mul_table));
  const int gap = 1 << kSigmaShift;
  int mul = -32768;
  for (int sigma = kMinSigma; sigma <= kMaxSigma; sigma += gap) {
    for (; mul < 0; ++mul) {
      if (WeightFromBits(sigma >> kSigmaShift, mul) > 0.5f) break;
    }
    t->mul_table[sigma] = mul;
  }
  for (int sigma = kMinSigma; sigma < kMaxSigma; sigma += gap) {
    const float mul_step =
        (t->mul_table[sigma + gap] - t->mul_table[sigma]) / float(gap);
    for (int i = 1; i < gap; ++i) {
      t->mul_table[sigma + i] =
          static_cast<int>(t->mul_table[sigma] + i * mul_step);
    }
  }
}

void GaborishWeights(int strength, float w[3]) {
  const float wu0 = 1.0f;
  float wu1, wu2;
  switch (strength) {
    case 1:  // k500
      wu1 = static_cast<float>(0.056007960760453189);
      wu2 = static_cast<float>(0.045899074552453879);
      break;
    case 2: {  // k750
      const float x = 0.75;
      wu1 = static_cast<float>(x * 0.13959942428275746);
      wu2 = static_cast<float>(x * 0.074240717189152386);
      break;
    }
    case 3: {  // k875
      const float x = 0.875;
      wu1 = static_cast<float>(x * 0.11501538179658321);
      wu2 = static_cast<float>(x * 0.089979079587015454);
      break;
    }
    default:  // k1000
      wu1 = static_cast<float>(0.11501538179658321);
      wu2 = static_cast<float>(0.089979079587015454);
      break;
  }
  const float mul = 1.0 / (wu0 + 4 * (wu1 + wu2));
  w[0] = wu0 * mul;
  w[1] = wu1 * mul;
  w[2] = wu2 * mul;
}

void LfBlurWeights(float w[3]) {
  w[0] = 0.41459272584128337f;
  w[1] = 0.25489157325704559f;
  w[2] = 0.046449679523692139f;
}

void Ub4Kernel(float k[3][4]) {
  const float sigma = 2.0007879236394901f;
  const int radius = 4;
  float kernel[2 * radius + 1];
  const float scaler = -1.0 / (2 * sigma * sigma);
  double sum = 0.0;
  for (int i = -radius; i <= radius; ++i) {
    const float val = expf(scaler * i * i);
    kernel[i + radius] = val;
    sum += val;
  }
  for (int i = 0; i < 2 * radius + 1; ++i) kernel[i] /= sum;
  for (int j = 0; j < 3; ++j) {
    for (int q = 0; q < 4; ++q) k[j][q] = 0.0f;
  }
  for (int q = 0; q < 4; ++q) {
    const int split0 = 4 - q;
    const int split1 = 8 - q;
    for (int j = 0; j < split0; ++j) k[0][q] += kernel[j];
    for (int j = split0; j < split1; ++j) k[1][q] += kernel[j];
    for (int j = split1; j < 2 * radius + 1; ++j) k[2][q] += kernel[j];
  }
}

// Transform scale tables, laid out as consts->tscale[log2idx][y][x] and
// ascale likewise; consumed through FillScaleTables.
void FillScaleTables(float* tscale, float* ascale) {
  for (int i = 0; i < 5; ++i) {
    const int n = 2 << i;
    for (int y = 0; y < 8; ++y) {
      for (int x = 0; x < 8; ++x) {
        const bool in = x < n && y < n;
        tscale[(i * 8 + y) * 8 + x] = in ? TotalScale(n, x, y) : 0.0f;
        ascale[(i * 8 + y) * 8 + x] = in ? InvTotalScale(n, x, y) : 0.0f;
      }
    }
  }
}

// D_k = sum_x b_k(x)^2 for b_k(x) = w_k cos((2x + 1) k pi / 2N), w_0 = 1,
// w_k = 1 / cos(k pi / 16) for N = 8 (the AAN scaling of the 8-point IDCT)
// and 2 cos(k pi / 2N) otherwise: N for k = 0, N/2 w_k^2 for k > 0.
void AdjointNorms(float* d) {
  for (int n = 4; n <= 32; n *= 2) {
    float* o = d + AdjNormOffset(n);
    o[0] = float(n);
    for (int k = 1; k < n; ++k) {
      const double w = n == 8 ? 1.0 / std::cos(k * M_PI / 16)
                              : 2.0 * std::cos(k * M_PI / (2 * n));
      o[k] = float(0.5 * n * w * w);
    }
  }
}

namespace {

// ApproxCubeRoot (pik/approx_cube_root.h): initial guess 0x2a50f200 + ix / 3
// on the float's bits, then two Newton steps x = 1/3 (2 x + y / x^2).
float CubeRootApprox(float y) {
  int32_t ix;
  memcpy(&ix, &y, 4);
  ix = 0x2a50f200 + ix / 3;
  float x;
  memcpy(&x, &ix, 4);
  const float kOneThird = 1.0f / 3.0f;
  for (int step = 0; step < 2; ++step) x = kOneThird * (2.0f * x + y / (x * x));
  return x;
}

// y ~= b * x + a over `size` points (GrayXyb::LinearRegression,
// pik/compressed_image.cc:164-193, its running-moment form).
void LinearRegression(const float* x, const float* y, size_t size, double* a,
                      double* b) {
  double mx = 0, my = 0;
  double mx2 = 0, my2 = 0;
  double mxy = 0;
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
  const double sx = std::sqrt(mx2 / (size - 1));
  const double sy = std::sqrt(my2 / (size - 1));
  const double sumxy = mxy * size + my * mx * size;
  const double r = (sumxy - size * mx * my) / ((size - 1.0) * sx * sy);
  *b = r * sy / sx;
  *a = my - *b * mx;
}

}  // namespace

void LinearToXybApprox(float r, float g, float b, float* x, float* y,
                       float* z) {
  // OpsinAbsorbance (pik/opsin_image.h:23-30), the matrix and bias of
  // pik/opsin_params.h:21-41 rounded from double.
  const float kScale = 255.0f;
  const float mix[9] = {
      float(0.29956550340058319 / kScale), float(0.63373087833825936 / kScale),
      float(0.077705617820981968 / kScale), float(0.22158691104574774 / kScale),
      float(0.68491388044116142 / kScale), float(0.096254234043612538 / kScale),
      float(0.20062661225219422 / kScale), float(0.070366199217588729 / kScale),
      float(0.5571760754215358 / kScale)};
  const float bias[3] = {float(0.26786006338144885 / kScale),
                         float(0.24494032763907073 / kScale),
                         float(0.14255999980363571 / kScale)};
  float mixed[3];
  for (int c = 0; c < 3; ++c) {
    const float v = mix[3 * c] * r + mix[3 * c + 1] * g + mix[3 * c + 2] * b +
                    bias[c];
    // std::max(0.0f, v): (0 < v) ? v : 0 (so -0 and NaN give +0)
    mixed[c] = CubeRootApprox(0.0f < v ? v : 0.0f);
  }
  // LinearXybTransform with kScaleR = kScaleG = 1 (pik/opsin_image.cc:25-30)
  *x = (1.0f * mixed[0] - 1.0f * mixed[1]) * 0.5f;
  *y = (1.0f * mixed[0] + 1.0f * mixed[1]) * 0.5f;
  *z = mixed[2];
}

void ComputeGrayXyb(GrayXybTable* t) {
  const int kN = 1024;
  std::vector<float> x(kN), y(kN), z(kN);
  for (int i = 0; i < kN; i++) {
    const float gray = (float)(256.0f * i / kN);
    LinearToXybApprox(gray, gray, gray, &x[i], &y[i], &z[i]);
  }
  const float mn = y[0], mx = y[kN - 1];
  int m = 0;
  int border[kGrayPieces + 1];
  for (int i = 0; i < kN; i++) {
    if (y[i] >= y[0] + (mx - mn) * m / kGrayPieces) {
      border[m] = i;
      m++;
    }
  }
  border[kGrayPieces] = kN;
  t->sub = mn;
  t->mul = 1.0 / (mx - mn);
  for (int i = 0; i < kGrayPieces; i++) {
    LinearRegression(y.data() + border[i], x.data() + border[i],
                     border[i + 1] - border[i], &t->x_const[i], &t->x_slope[i]);
    LinearRegression(y.data() + border[i], z.data() + border[i],
                     border[i + 1] - border[i], &t->b_const[i], &t->b_slope[i]);
  }
}

}  // namespace pikgpu
