// One-dimensional DCT-II / DCT-III butterflies of the transposed-scaled
// transforms, for host and device.
//
// The reference computes its block transforms with two fast algorithms
// (pik/dct.h:25-40, pik/dct_simd_any.h):
//  - 8 points: the Arai-Agui-Nakajima flow graph (JPEG: Still Image Data
//    Compression Standard, section 4.3.5), used for DCT8 blocks;
//  - 4, 16, 32 points (and the 8-point halves inside 16/32): the recursive
//    algorithm of Vashkevich and Petrovsky, "A low multiplicative complexity
//    fast recursive DCT-2 algorithm" (2012).
// The EPF guide is an 8-bit quantisation of the reconstructed image, so the
// last float bit of the IDCT decides some guide bytes. To reproduce the
// reference's bytes the transforms here follow the same flow graphs: every
// float operation has the same operands (up to the commutativity of + and *).
// They are written as generic recursions; the constants are 2cos(theta)
// (forward) and 0.5/cos(theta) (inverse) rounded from double.
#pragma once

#include <math.h>

#include "dct_consts.h"

// Everything is forced inline with fully unrolled loops and compile-time
// angles: every array index is then a constant and the arrays live in
// registers (a runtime index puts an array in per-lane scratch memory).
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define PIKGPU_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define PIKGPU_HD inline __attribute__((always_inline))
#endif

namespace pikgpu {
namespace dct1d {

// Angles are kept as integers J in units of pi / (2n) for an n-point
// transform (template arguments); the constants 2cos(J pi/2n) and
// 0.5/cos(J pi/2n) come from the tables of dct_consts.h.
template <int n>
PIKGPU_HD float FwdConst(int J) {
  return n == 4 ? kFwdConst4[J]
                : (n == 8 ? kFwdConst8[J]
                          : (n == 16 ? kFwdConst16[J] : kFwdConst32[J]));
}
template <int n>
PIKGPU_HD float InvConst(int J) {
  return n == 4 ? kInvConst4[J]
                : (n == 8 ? kInvConst8[J]
                          : (n == 16 ? kInvConst16[J] : kInvConst32[J]));
}

// ---------------------------------------------------------------------------
// Forward (DCT-II, unnormalised) in the Vashkevich-Petrovsky form.
//
// Odd part: node of m inputs at angle J computes a_k = u_k - u_{m-1-k},
// b_k = u_{m/2+k} * 2cos(J pi/2n), p = a + b, q = a - b; a node of 2 inputs
// writes y[J/2] = p, y[n - J/2] = q, larger nodes recurse on p at J/2 and on
// q at (2n - J)/2.
template <int n, int m, int J>
struct VpOdd {
  PIKGPU_HD static void Run(const float* u, float* y) {
    float p[m / 2], q[m / 2];
    const float c = FwdConst<n>(J);
#pragma unroll
    for (int k = 0; k < m / 2; ++k) {
      const float a = u[k] - u[m - 1 - k];
      const float b = u[m / 2 + k] * c;
      p[k] = a + b;
      q[k] = a - b;
    }
    VpOdd<n, m / 2, J / 2>::Run(p, y);
    VpOdd<n, m / 2, (2 * n - J) / 2>::Run(q, y);
  }
};
template <int n, int J>
struct VpOdd<n, 2, J> {
  PIKGPU_HD static void Run(const float* u, float* y) {
    const float a = u[0] - u[1];
    const float b = u[1] * FwdConst<n>(J);
    y[J / 2] = a + b;
    y[n - J / 2] = a - b;
  }
};

template <int n>
struct VpDct {
  PIKGPU_HD static void Run(const float* x, float* y) {
    float s[n / 2], d[n / 2], e[n / 2];
#pragma unroll
    for (int k = 0; k < n / 2; ++k) {
      s[k] = x[k] + x[n - 1 - k];
      d[k] = x[k] - x[n - 1 - k];
    }
    VpDct<n / 2>::Run(s, e);
#pragma unroll
    for (int k = 0; k < n / 2; ++k) y[2 * k] = e[k];
    VpOdd<n, n / 2, n / 2>::Run(d, y);
  }
};
template <>
struct VpDct<2> {
  PIKGPU_HD static void Run(const float* x, float* y) {
    y[0] = x[0] + x[1];
    y[1] = x[0] - x[1];
  }
};

// ---------------------------------------------------------------------------
// Inverse (DCT-III) in the Vashkevich-Petrovsky form.
//
// Odd part: groups are merged pairwise; merging G and H (h values each) with
// c = 0.5/cos(J pi/2n) gives S_k = G_k + H_k, D_k = (G_k - H_k) c and the
// group (S_k + D_{h-1-k})_k followed by (D_k)_k. Leaves merge the single
// inputs x_j and x_{n-j}, j = J/2.
template <int n, int size, int J>
struct VpInvOdd {
  PIKGPU_HD static void Run(const float* x, float* out) {
    float g[size / 2], h[size / 2];
    VpInvOdd<n, size / 2, J / 2>::Run(x, g);
    VpInvOdd<n, size / 2, (2 * n - J) / 2>::Run(x, h);
    const float c = InvConst<n>(J);
    float s[size / 2], d[size / 2];
#pragma unroll
    for (int k = 0; k < size / 2; ++k) {
      s[k] = g[k] + h[k];
      d[k] = (g[k] - h[k]) * c;
    }
#pragma unroll
    for (int k = 0; k < size / 2; ++k) {
      out[k] = s[k] + d[size / 2 - 1 - k];
      out[size / 2 + k] = d[k];
    }
  }
};
template <int n, int J>
struct VpInvOdd<n, 2, J> {
  PIKGPU_HD static void Run(const float* x, float* out) {
    constexpr int j = J / 2;
    const float s = x[j] + x[n - j];
    const float d = (x[j] - x[n - j]) * InvConst<n>(J);
    out[0] = s + d;
    out[1] = d;
  }
};

template <int n>
struct VpIdct {
  PIKGPU_HD static void Run(const float* x, float* y) {
    float ev[n / 2], e[n / 2], o[n / 2];
#pragma unroll
    for (int k = 0; k < n / 2; ++k) ev[k] = x[2 * k];
    VpIdct<n / 2>::Run(ev, e);
    VpInvOdd<n, n / 2, n / 2>::Run(x, o);
#pragma unroll
    for (int k = 0; k < n / 2; ++k) {
      y[k] = e[k] + o[k];
      y[n - 1 - k] = e[k] - o[k];
    }
  }
};
template <>
struct VpIdct<2> {
  PIKGPU_HD static void Run(const float* x, float* y) {
    y[0] = x[0] + x[1];
    y[1] = x[0] - x[1];
  }
};

// ---------------------------------------------------------------------------
// 8-point Arai-Agui-Nakajima flow graphs (scaled; the scale factors live in
// the dequantisation tables).
PIKGPU_HD void AanDct8(const float* i, float* o) {
  const float c1 = 0.707106781186548f;  // 1 / sqrt(2)
  const float c2 = 0.382683432365090f;  // cos(3 pi / 8)
  const float c3 = 1.30656296487638f;   // 1 / (2 cos(3 pi / 8))
  const float c4 = 0.541196100146197f;  // sqrt(2) cos(3 pi / 8)
  // stage 1: sums and differences of mirrored inputs
  const float s07 = i[0] + i[7], d07 = i[0] - i[7];
  const float s34 = i[3] + i[4], d34 = i[3] - i[4];
  const float s25 = i[2] + i[5], d25 = i[2] - i[5];
  const float s16 = i[1] + i[6], d16 = i[1] - i[6];
  // even part
  const float e0 = s07 + s34, e1 = s07 - s34;
  const float e2 = s16 + s25, e3 = s16 - s25;
  const float r = (e3 + e1) * c1;
  // odd part
  const float o12 = d16 + d25;
  const float o01 = d07 + d16;
  const float o23 = d25 + d34;
  const float z = (o23 - o01) * c2;
  const float m = o12 * c1;
  const float a = d07 + m, b = d07 - m;
  const float u = c3 * o01 + z;
  const float v = c4 * o23 + z;
  o[0] = e0 + e2;
  o[1] = a + u;
  o[2] = e1 + r;
  o[3] = b - v;
  o[4] = e0 - e2;
  o[5] = b + v;
  o[6] = e1 - r;
  o[7] = a - u;
}

PIKGPU_HD void AanIdct8(const float* i, float* o) {
  const float c1 = 1.41421356237310f;  // sqrt(2)
  const float c2 = 2.61312592975275f;  // 1 / cos(3 pi / 8)
  const float c3 = 0.76536686473018f;  // 2 cos(3 pi / 8)
  const float c4 = 1.08239220029239f;  // 2 sqrt(2) cos(3 pi / 8)
  const float p04 = i[0] + i[4], m04 = i[0] - i[4];
  const float p62 = i[6] + i[2], m62 = i[6] - i[2];
  const float p71 = i[7] + i[1], m71 = i[7] - i[1];
  const float p53 = i[5] + i[3], m53 = i[5] - i[3];
  const float q0 = p71 + p53, q1 = p71 - p53;
  const float e0 = p04 + p62, e1 = p04 - p62;
  const float w = (m53 - m71) * c3;
  const float t = c1 * m62 + p62;
  const float f0 = m04 - t, f1 = m04 + t;
  const float g0 = c2 * m71 + w;
  const float g1 = c4 * m53 + w;
  const float h0 = q0 + g0;
  const float h1 = c1 * q1 + h0;
  const float h2 = g1 - h1;
  o[0] = e0 + q0;
  o[1] = f0 - h0;
  o[2] = f1 + h1;
  o[3] = e1 + h2;
  o[4] = e1 - h2;
  o[5] = f1 - h1;
  o[6] = f0 + h0;
  o[7] = e0 - q0;
}

// ---------------------------------------------------------------------------
// Dispatch: forward and inverse 1-D transforms of n points as used for
// standalone n x n blocks.
template <int n>
PIKGPU_HD void Dct(const float* x, float* y) {
  VpDct<n>::Run(x, y);
}
template <>
PIKGPU_HD void Dct<8>(const float* x, float* y) {
  AanDct8(x, y);
}
template <int n>
PIKGPU_HD void Idct(const float* x, float* y) {
  VpIdct<n>::Run(x, y);
}
template <>
PIKGPU_HD void Idct<8>(const float* x, float* y) {
  AanIdct8(x, y);
}

// Two-dimensional transposed-scaled transforms of one n x n block with the
// reference's pass order (columns, transpose, columns; pik/dct.h:370-507).
// coef is row-major n x n with the horizontal frequency as row index.
template <int n>
PIKGPU_HD void Idct2D(const float* coef, int cstride, float* pix,
                      int pstride) {
  float t[n * n];
#pragma unroll
  for (int c = 0; c < n; ++c) {
    float v[n], w[n];
#pragma unroll
    for (int r = 0; r < n; ++r) v[r] = coef[r * cstride + c];
    Idct<n>(v, w);
#pragma unroll
    for (int x = 0; x < n; ++x) t[c * n + x] = w[x];
  }
#pragma unroll
  for (int x = 0; x < n; ++x) {
    float v[n], w[n];
#pragma unroll
    for (int c = 0; c < n; ++c) v[c] = t[c * n + x];
    Idct<n>(v, w);
#pragma unroll
    for (int y = 0; y < n; ++y) pix[y * pstride + x] = w[y];
  }
}

template <int n>
PIKGPU_HD void Dct2D(const float* pix, int pstride, float* coef,
                     int cstride) {
  float t[n * n];
#pragma unroll
  for (int x = 0; x < n; ++x) {
    float v[n], w[n];
#pragma unroll
    for (int y = 0; y < n; ++y) v[y] = pix[y * pstride + x];
    Dct<n>(v, w);
#pragma unroll
    for (int k = 0; k < n; ++k) t[x * n + k] = w[k];
  }
  const float scale = 1.0f / (n * n);
#pragma unroll
  for (int k = 0; k < n; ++k) {
    float v[n], w[n];
#pragma unroll
    for (int x = 0; x < n; ++x) v[x] = t[x * n + k];
    Dct<n>(v, w);
#pragma unroll
    for (int j = 0; j < n; ++j) coef[j * cstride + k] = w[j] * scale;
  }
}

}  // namespace dct1d
}  // namespace pikgpu
