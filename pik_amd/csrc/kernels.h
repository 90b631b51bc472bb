// Device helpers and kernel declarations of the reconstruction (gfx950).
//
// All arithmetic that feeds an integer decision later on (the EPF guide
// bytes and sigma) keeps the reference's operation order with separate
// multiplies and adds; the library is compiled with -ffp-contract=off.
#pragma once

#include "device.h"
#include "xform.h"

namespace pikgpu {

// ---------------------------------------------------------------------------
// Transposed-scaled transforms.
//
// Pixels P (N x N, [y][x]) and coefficients F (N x N, row-major) relate by
// P = B^T F^T B with the float basis B of ComputeBasis: the coefficient row
// index is the horizontal frequency (pik/dct.h:351-507).

__device__ inline int BasisOffsetD(int n) {
  return n == 2 ? 0 : (n == 4 ? 8 : (n == 8 ? 40 : (n == 16 ? 168 : 680)));
}
__device__ inline const float* BasisB(const DeviceConsts* K, int n) {
  return K->basis + BasisOffsetD(n);
}
__device__ inline const float* BasisInv(const DeviceConsts* K, int n) {
  return K->basis + BasisOffsetD(n) + n * n;
}

// In-thread ts-IDCT / ts-DCT (with the 1/N^2 scaling of ScaleToBlock) of an
// n x n block, n = 2, 4, 8, as the reinterpreting transforms use them
// (pik/ac_strategy.cc:272-315); n == 2 is the reference's explicit 2-point
// form (pik/dct.h:478-504).
__device__ inline void SmallIdct(const DeviceConsts* K, const float* in,
                                 int in_stride, float* out, int out_stride,
                                 int n) {
  if (n == 2) {
    const float a00 = in[0], a01 = in[1], a10 = in[in_stride],
                a11 = in[in_stride + 1];
    out[0] = a00 + a01 + a10 + a11;
    out[1] = a00 + a01 - a10 - a11;
    out[out_stride] = a00 - a01 + a10 - a11;
    out[out_stride + 1] = a00 - a01 - a10 + a11;
  } else if (n == 4) {
    dct1d::Idct2D<4>(in, in_stride, out, out_stride);
  } else {
    dct1d::Idct2D<8>(in, in_stride, out, out_stride);
  }
}

__device__ inline void SmallDct(const DeviceConsts* K, const float* in,
                                int in_stride, float* out, int out_stride,
                                int n) {
  if (n == 2) {
    const float a00 = in[0], a01 = in[1], a10 = in[in_stride],
                a11 = in[in_stride + 1];
    out[0] = (a00 + a01 + a10 + a11) * 0.25f;
    out[1] = (a00 + a01 - a10 - a11) * 0.25f;
    out[out_stride] = (a00 - a01 + a10 - a11) * 0.25f;
    out[out_stride + 1] = (a00 - a01 - a10 + a11) * 0.25f;
  } else if (n == 4) {
    dct1d::Dct2D<4>(in, in_stride, out, out_stride);
  } else {
    dct1d::Dct2D<8>(in, in_stride, out, out_stride);
  }
}

// 3x3 symmetric convolution of one plane at (x, y) with mirrored borders,
// in the operation order of strategy::Symmetric3 (pik/convolve.h:605-621).
__device__ inline float Symmetric3(const float* in, int pitch, int W, int H,
                                   int x, int y, float w0, float w1,
                                   float w2) {
  const int xm = Mirror(x - 1, W), xp = Mirror(x + 1, W);
  const int ym = Mirror(y - 1, H), yp = Mirror(y + 1, H);
  const float* rt = in + ym * pitch;
  const float* rm = in + y * pitch;
  const float* rb = in + yp * pitch;
  const float sum_tb = rt[x] + rb[x];
  const float mul0 = rm[x] * w0;
  const float sum_lr = rm[xm] + rm[xp];
  const float x1 = sum_tb + sum_lr;
  const float mul1 = x1 * w1 + mul0;
  const float sum_t2 = rt[xm] + rt[xp];
  const float sum_b2 = rb[xm] + rb[xp];
  const float x2 = sum_t2 + sum_b2;
  return x2 * w2 + mul1;
}

// AdjustQuantBias, vector form (pik/quantizer.h:61-88).
__device__ inline float AdjustQuantBias(int c, float q) {
  const float one_bias = c == 0 ? 0.05465007330715401f
                                : (c == 1 ? 0.07005449891748593f
                                          : 0.049935103337343655f);
  const float aq = fabsf(q);
  if (aq >= 1.5f) return q - 0.145f / q;
  if (aq >= 0.5f) return copysignf(1 - one_bias, q);
  return q;
}

// Dequantised (and for X/B colour-correlated) AC coefficient at frame-wide
// storage position (row = block row, col = block x * 64 + k) of plane c;
// `yval` is the dequantised Y value at the same position
// (Dequant::DoAC, pik/compressed_image.cc:621-693).
__device__ inline float DequantCoef(const FrameParams& P, const FrameBuffers& B,
                                    int c, int row, int col, float yval) {
  const int bxa = col >> 6, k = col & 63;
  const int blk = row * P.bx + bxa;
  const uint8_t a = B.acs[blk];
  const int raw = a >> 4, idx = a & 15;
  const int qf = B.qf[blk];
  const int tile = (row >> 3) * P.tx + (bxa >> 3);
  const int qt = B.dqmap[B.dcf[tile] * 256 + qf - 1];
  const int kind = AcsQuantKind(raw);
  const float m = B.table[B.offsets[(qt * kNumQuantKinds + kind) * 3 + c] +
                          idx * 64 + k];
  const float scaled = qf == 0 ? 1E10f : P.inv_global_scale / float(qf);
  const float mul = m * scaled;
  const float q = float(B.ac[(size_t(c) * P.by + row) * (P.bx * 64) + col]);
  const float d = AdjustQuantBias(c, q) * mul;
  if (c == 1) return d;
  float f;
  if (c == 0) {
    f = float(B.ytox[tile] - 128) * 0.00390625f;
  } else {
    f = float(B.ytob[tile]) * 0.0078125f;
  }
  return f * yval + d;
}

// ---------------------------------------------------------------------------
// Cooperative transforms of one channel of a 32x32 tile (256 threads).
// `acs_t` holds the 16 strategy bytes of the tile's 8x8 blocks (0xFF outside
// the image). Coefficients live in the reference's storage layout: 4 block
// rows x 256 floats. The 2-D transforms run as the reference's two column
// passes with a transpose in between (pik/dct.h:370-507): every line is a
// 1-D butterfly of dct1d.h computed by one thread. Thread t works for the
// 8x8 block t / 16 of the tile; only threads of the first block of a
// transform have lines to do (16 per block: a DCT8 has 8 lines, a DCT4X4 4
// lines in each of its 4 sub-blocks, DCT16 16 and DCT32 32, two per thread).

struct TileXform {
  int raw, s, N;   // strategy, covered blocks, transform size
  int oby, obx;    // origin block of this thread's transform in the tile
  int lane;        // thread index within its 8x8 block (0..15)
  bool first;      // this thread's block is the transform's first block
  bool valid;      // block inside the image
  // pixel-quad ownership for elementwise work: row r, columns c0..c0+3
  int r, c0;
};

__device__ inline TileXform MakeXform(const uint8_t* acs_t, int tid) {
  TileXform X;
  X.r = tid >> 3;
  X.c0 = (tid & 7) << 2;
  const int blk = tid >> 4;
  X.lane = tid & 15;
  const int lby = blk >> 2, lbx = blk & 3;
  const uint8_t a = acs_t[blk];
  X.valid = a != 0xFF;
  X.raw = X.valid ? (a >> 4) : 0;
  const int idx = X.valid ? (a & 15) : 0;
  X.s = AcsCovered(X.raw);
  X.N = 8 * X.s;
  X.oby = lby - idx / X.s;
  X.obx = lbx - idx % X.s;
  X.first = X.valid && idx == 0;
  return X;
}

// Coefficient storage index of F[i][j] (row-major N x N, flat i*N + j) of a
// transform with origin block (oby, obx) covering s x s blocks.
__device__ inline int CoefIdx(int oby, int obx, int s, int flat) {
  return (oby + flat / (64 * s)) * 256 + obx * 64 + flat % (64 * s);
}

template <int N>
__device__ inline void InvLine1(const float* coef, float* tmp, int oby,
                                int obx, int c) {
  constexpr int s = N / 8;
  float v[N], w[N];
  for (int r = 0; r < N; ++r) v[r] = coef[CoefIdx(oby, obx, s, r * N + c)];
  dct1d::Idct<N>(v, w);
  for (int x = 0; x < N; ++x) tmp[(oby * 8 + c) * 32 + obx * 8 + x] = w[x];
}
template <int N>
__device__ inline void InvLine2(const float* tmp, float* pix, int oby,
                                int obx, int x) {
  float v[N], w[N];
  for (int c = 0; c < N; ++c) v[c] = tmp[(oby * 8 + c) * 32 + obx * 8 + x];
  dct1d::Idct<N>(v, w);
  for (int y = 0; y < N; ++y) pix[(oby * 8 + y) * 32 + obx * 8 + x] = w[y];
}
template <int N>
__device__ inline void FwdLine1(const float* pix, float* tmp, int oby,
                                int obx, int x) {
  float v[N], w[N];
  for (int y = 0; y < N; ++y) v[y] = pix[(oby * 8 + y) * 32 + obx * 8 + x];
  dct1d::Dct<N>(v, w);
  for (int k = 0; k < N; ++k) tmp[(oby * 8 + x) * 32 + obx * 8 + k] = w[k];
}
template <int N>
__device__ inline void FwdLine2(const float* tmp, float* coef, int oby,
                                int obx, int k) {
  constexpr int s = N / 8;
  float v[N], w[N];
  for (int x = 0; x < N; ++x) v[x] = tmp[(oby * 8 + x) * 32 + obx * 8 + k];
  dct1d::Dct<N>(v, w);
  const float scale = 1.0f / (N * N);
  for (int j = 0; j < N; ++j) {
    coef[CoefIdx(oby, obx, s, j * N + k)] = w[j] * scale;
  }
}

// coef (4 x 256) -> pix (32 x 32). All 256 threads must call.
__device__ inline void TileInverse(const DeviceConsts* K, const TileXform& X,
                                   const float* coef, float* pix, float* tmp) {
  if (X.first) {
    const int l = X.lane;
    if (X.raw == kDCT || X.raw == kDCTNoHF) {
      if (l < 8) InvLine1<8>(coef, tmp, X.oby, X.obx, l);
    } else if (X.raw == kDCT16) {
      InvLine1<16>(coef, tmp, X.oby, X.obx, l);
    } else if (X.raw == kDCT32) {
      InvLine1<32>(coef, tmp, X.oby, X.obx, l);
      InvLine1<32>(coef, tmp, X.oby, X.obx, l + 16);
    } else if (AcsIsDCT4(X.raw)) {
      // sub-block (sy, sx), column c of its 4x4 coefficient matrix G with
      // G[iy][ix] = coef[(sy + 2 iy) * 8 + sx + 2 ix], G[0][0] from the
      // Hadamard of the four DCs (pik/ac_strategy.cc:762-790)
      const float* cb = coef + X.oby * 256 + X.obx * 64;
      const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, c = l & 3;
      const float b00 = cb[0], b01 = cb[1], b10 = cb[8], b11 = cb[9];
      float dc;
      if (sub == 0) dc = b00 + b01 + b10 + b11;
      else if (sub == 1) dc = b00 + b01 - b10 - b11;
      else if (sub == 2) dc = b00 - b01 + b10 - b11;
      else dc = b00 - b01 - b10 + b11;
      float v[4], w[4];
      for (int r = 0; r < 4; ++r) {
        v[r] = (r == 0 && c == 0) ? dc : cb[(sy + r * 2) * 8 + sx + c * 2];
      }
      dct1d::Idct<4>(v, w);
      for (int x = 0; x < 4; ++x) {
        tmp[(X.oby * 8 + sy * 4 + c) * 32 + X.obx * 8 + sx * 4 + x] = w[x];
      }
    }
  }
  __syncthreads();
  if (X.first) {
    const int l = X.lane;
    if (X.raw == kDCT || X.raw == kDCTNoHF) {
      if (l < 8) InvLine2<8>(tmp, pix, X.oby, X.obx, l);
    } else if (X.raw == kDCT16) {
      InvLine2<16>(tmp, pix, X.oby, X.obx, l);
    } else if (X.raw == kDCT32) {
      InvLine2<32>(tmp, pix, X.oby, X.obx, l);
      InvLine2<32>(tmp, pix, X.oby, X.obx, l + 16);
    } else if (AcsIsDCT4(X.raw)) {
      const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, x = l & 3;
      float v[4], w[4];
      for (int c = 0; c < 4; ++c) {
        v[c] = tmp[(X.oby * 8 + sy * 4 + c) * 32 + X.obx * 8 + sx * 4 + x];
      }
      dct1d::Idct<4>(v, w);
      for (int y = 0; y < 4; ++y) {
        pix[(X.oby * 8 + sy * 4 + y) * 32 + X.obx * 8 + sx * 4 + x] = w[y];
      }
    } else if (l == 0) {
      DirectInverse8(X.raw, coef + X.oby * 256 + X.obx * 64,
                     pix + X.oby * 8 * 32 + X.obx * 8, 32);
    }
  }
  __syncthreads();
}

// pix (32 x 32) -> coef (4 x 256). `hf_only` restricts to strategies with
// HF prediction (others keep their coef entries untouched).
__device__ inline void TileForward(const DeviceConsts* K, const TileXform& X,
                                   const float* pix, float* coef, float* tmp,
                                   bool hf_only) {
  const bool active = X.first && (!hf_only || AcsPredictHF(X.raw));
  const int l = X.lane;
  if (active) {
    if (X.raw == kDCT || X.raw == kDCTNoHF) {
      if (l < 8) FwdLine1<8>(pix, tmp, X.oby, X.obx, l);
    } else if (X.raw == kDCT16) {
      FwdLine1<16>(pix, tmp, X.oby, X.obx, l);
    } else if (X.raw == kDCT32) {
      FwdLine1<32>(pix, tmp, X.oby, X.obx, l);
      FwdLine1<32>(pix, tmp, X.oby, X.obx, l + 16);
    } else if (AcsIsDCT4(X.raw)) {
      const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, x = l & 3;
      float v[4], w[4];
      for (int y = 0; y < 4; ++y) {
        v[y] = pix[(X.oby * 8 + sy * 4 + y) * 32 + X.obx * 8 + sx * 4 + x];
      }
      dct1d::Dct<4>(v, w);
      for (int k = 0; k < 4; ++k) {
        tmp[(X.oby * 8 + sy * 4 + x) * 32 + X.obx * 8 + sx * 4 + k] = w[k];
      }
    }
  }
  __syncthreads();
  if (active) {
    if (X.raw == kDCT || X.raw == kDCTNoHF) {
      if (l < 8) FwdLine2<8>(tmp, coef, X.oby, X.obx, l);
    } else if (X.raw == kDCT16) {
      FwdLine2<16>(tmp, coef, X.oby, X.obx, l);
    } else if (X.raw == kDCT32) {
      FwdLine2<32>(tmp, coef, X.oby, X.obx, l);
      FwdLine2<32>(tmp, coef, X.oby, X.obx, l + 16);
    } else if (AcsIsDCT4(X.raw)) {
      const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, k = l & 3;
      float v[4], w[4];
      for (int x = 0; x < 4; ++x) {
        v[x] = tmp[(X.oby * 8 + sy * 4 + x) * 32 + X.obx * 8 + sx * 4 + k];
      }
      dct1d::Dct<4>(v, w);
      float* cb = coef + X.oby * 256 + X.obx * 64;
      for (int j = 0; j < 4; ++j) cb[(sy + j * 2) * 8 + sx + k * 2] = w[j] * (1.0f / 16);
    } else if (l == 0) {
      DirectForward8(X.raw, pix + X.oby * 8 * 32 + X.obx * 8, 32,
                     coef + X.oby * 256 + X.obx * 64);
    }
  }
  __syncthreads();
  if (active && AcsIsDCT4(X.raw) && l == 0) {
    float* cb = coef + X.oby * 256 + X.obx * 64;
    const float b00 = cb[0], b01 = cb[1], b10 = cb[8], b11 = cb[9];
    cb[0] = (b00 + b01 + b10 + b11) * 0.25f;
    cb[1] = (b00 + b01 - b10 - b11) * 0.25f;
    cb[8] = (b00 - b01 + b10 - b11) * 0.25f;
    cb[9] = (b00 - b01 - b10 + b11) * 0.25f;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Kernels.
__global__ void KDcDequant(FrameParams P, FrameBuffers B);
__global__ void KComputeLlf(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KDc2x2FromLlf(FrameParams P, FrameBuffers B,
                              const DeviceConsts* K, int to_pred2x2);
__global__ void KLfBlur(FrameParams P, FrameBuffers B);
__global__ void KLfFromDc2x2(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KZeroPred(FrameParams P, FrameBuffers B);
__global__ void KPred2x2Border(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KPred2x2Interior(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KBlocks(FrameParams P, FrameBuffers B, const DeviceConsts* K);

// Generic image description for the filters.
struct Img {
  const float* p;  // 3 planes
  size_t plane;    // floats per plane
  int pitch;       // floats per row
  int W, H;        // logical size (mirror borders)
  int vw, vh;      // valid source region (replicate beyond, <= W, H)
};

__global__ void KMinMaxPartial(Img in, int gaborish, FrameParams P,
                               float* partial);
__global__ void KMinMaxFinal(const float* partial, int n, float* mm);
__global__ void KEpf(Img in, int gaborish, int dc_mode, FrameParams P,
                     FrameBuffers B, const DeviceConsts* K, const float* mm,
                     float* out, int out_pitch, size_t out_plane);
__global__ void KDots(FrameParams P, FrameBuffers B);
__global__ void KFinal(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KCopyPlanes(const float* src, float* dst, size_t n);

}  // namespace pikgpu
