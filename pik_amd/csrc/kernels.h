// Device helpers and kernel declarations of the reconstruction (gfx950).
//
// All arithmetic that feeds an integer decision later on (the EPF guide
// bytes and sigma) keeps the reference's operation order with separate
// multiplies and adds; the library is compiled with -ffp-contract=off.
#pragma once

#include "device.h"

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

// In-thread ts-IDCT of an n x n block (n <= 8). n == 2 uses the reference's
// explicit 2-point form (pik/dct.h:491-504).
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
    return;
  }
  const float* b = BasisB(K, n);
  float t[64];
  // t[c][x] = sum_r F[r][c] B[r][x]
  for (int c = 0; c < n; ++c) {
    for (int x = 0; x < n; ++x) {
      float acc = 0.0f;
      for (int r = 0; r < n; ++r) acc = fmaf(in[r * in_stride + c], b[r * n + x], acc);
      t[c * n + x] = acc;
    }
  }
  for (int y = 0; y < n; ++y) {
    for (int x = 0; x < n; ++x) {
      float acc = 0.0f;
      for (int c = 0; c < n; ++c) acc = fmaf(b[c * n + y], t[c * n + x], acc);
      out[y * out_stride + x] = acc;
    }
  }
}

// In-thread ts-DCT (with the 1/N^2 scaling of ScaleToBlock) of n x n.
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
    return;
  }
  const float* bi = BasisInv(K, n);
  float u[64];
  // u[y][r] = sum_x P[y][x] Binv[x][r]
  for (int y = 0; y < n; ++y) {
    for (int r = 0; r < n; ++r) {
      float acc = 0.0f;
      for (int x = 0; x < n; ++x) acc = fmaf(in[y * in_stride + x], bi[x * n + r], acc);
      u[y * n + r] = acc;
    }
  }
  // F[r][c] = sum_y Binv[y][c] u[y][r]
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) {
      float acc = 0.0f;
      for (int y = 0; y < n; ++y) acc = fmaf(bi[y * n + c], u[y * n + r], acc);
      out[r * out_stride + c] = acc;
    }
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
// Cooperative transforms of one channel of a 32x32 tile (256 threads, each
// owning a 4-pixel row segment). `acs_t` holds the 16 strategy bytes of the
// tile's 8x8 blocks (0xFF outside the image). Coefficients live in the
// reference's storage layout: 4 block rows x 256 floats.

struct TileXform {
  int r, c0;       // thread's pixel row and first column in the tile
  int kind;        // 0 none, 1 separable DCT N, 2 DCT4X4, 3 direct 8x8
  int raw, s, N;   // strategy, covered blocks, transform size
  int oby, obx;    // origin block of the transform in the tile
  int ly, lx;      // thread position relative to the origin (pixels)
  bool owner;      // owns an 8x8 direct-kind block
};

__device__ inline TileXform MakeXform(const uint8_t* acs_t, int tid) {
  TileXform X;
  X.r = tid >> 3;
  X.c0 = (tid & 7) << 2;
  const int lby = X.r >> 3, lbx = X.c0 >> 3;
  const uint8_t a = acs_t[lby * 4 + lbx];
  X.owner = false;
  if (a == 0xFF) {
    X.kind = 0;
    X.raw = 0; X.s = 1; X.N = 8; X.oby = lby; X.obx = lbx; X.ly = 0; X.lx = 0;
    return X;
  }
  X.raw = a >> 4;
  const int idx = a & 15;
  X.s = AcsCovered(X.raw);
  X.N = 8 * X.s;
  X.oby = lby - idx / X.s;
  X.obx = lbx - idx % X.s;
  X.ly = X.r - X.oby * 8;
  X.lx = X.c0 - X.obx * 8;
  if (AcsIsDCT(X.raw)) {
    X.kind = 1;
  } else if (AcsIsDCT4(X.raw)) {
    X.kind = 2;
  } else {
    X.kind = 3;
    X.owner = (X.r & 7) == 0 && (X.c0 & 7) == 0;
  }
  return X;
}

__device__ inline float& CoefRef(float* coef, const TileXform& X, int flat) {
  const int row = X.oby + flat / (64 * X.s);
  const int col = X.obx * 64 + flat % (64 * X.s);
  return coef[row * 256 + col];
}

// Direct 8x8 inverse transforms: DCT2X2 pyramid (pik/ac_strategy.cc:
// 791-803, 330-356), IDENTITY (pik/ac_strategy.cc:725-761), and LINES is not
// emitted by the encoder (pik/ac_strategy.cc:18) and decodes to zeros here.
__device__ inline void DirectInverse8(int raw, const float* c, float* p,
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
__device__ inline void DirectForward8(int raw, const float* p, int pstride,
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

// coef (4 x 256) -> pix (32 x 32). All 256 threads must call.
__device__ inline void TileInverse(const DeviceConsts* K, const TileXform& X,
                                   const float* coef, float* pix, float* tmp) {
  if (X.kind == 1) {
    const float* b = BasisB(K, X.N);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < X.N; ++q) {
      const int flat = q * X.N + X.ly;
      const float f = coef[(X.oby + flat / (64 * X.s)) * 256 + X.obx * 64 +
                           flat % (64 * X.s)];
      const float* br = b + q * X.N + X.lx;
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(f, br[j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) tmp[X.r * 32 + X.c0 + j] = acc[j];
  } else if (X.kind == 2) {
    const float* cb = coef + X.oby * 256 + X.obx * 64;
    const int sy = X.ly >> 2, sx = X.lx >> 2, l = X.ly & 3;
    const float b00 = cb[0], b01 = cb[1], b10 = cb[8], b11 = cb[9];
    float dc;
    const int d = sy * 2 + sx;
    if (d == 0) dc = b00 + b01 + b10 + b11;
    else if (d == 1) dc = b00 + b01 - b10 - b11;
    else if (d == 2) dc = b00 - b01 + b10 - b11;
    else dc = b00 - b01 - b10 + b11;
    const float* b = BasisB(K, 4);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < 4; ++q) {
      const float g = (q == 0 && l == 0) ? dc : cb[(sy + q * 2) * 8 + sx + l * 2];
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(g, b[q * 4 + j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) tmp[X.r * 32 + X.c0 + j] = acc[j];
  }
  __syncthreads();
  if (X.kind == 1) {
    const float* b = BasisB(K, X.N);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < X.N; ++q) {
      const float w = b[q * X.N + X.ly];
      const float* tr = tmp + (X.oby * 8 + q) * 32 + X.c0;
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(w, tr[j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) pix[X.r * 32 + X.c0 + j] = acc[j];
  } else if (X.kind == 2) {
    const float* b = BasisB(K, 4);
    const int sy = X.ly >> 2, l = X.ly & 3;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < 4; ++q) {
      const float w = b[q * 4 + l];
      const float* tr = tmp + (X.oby * 8 + sy * 4 + q) * 32 + X.c0;
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(w, tr[j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) pix[X.r * 32 + X.c0 + j] = acc[j];
  } else if (X.kind == 3 && X.owner) {
    DirectInverse8(X.raw, coef + X.oby * 256 + X.obx * 64,
                   pix + X.r * 32 + X.c0, 32);
  }
  __syncthreads();
}

// pix (32 x 32) -> coef (4 x 256). `hf_only` restricts to strategies with
// HF prediction (others keep their coef entries untouched).
__device__ inline void TileForward(const DeviceConsts* K, const TileXform& X,
                                   const float* pix, float* coef, float* tmp,
                                   bool hf_only) {
  const bool active = X.kind != 0 && (!hf_only || AcsPredictHF(X.raw));
  if (active && X.kind == 1) {
    const float* bi = BasisInv(K, X.N);
    const float* pr = pix + X.r * 32 + X.obx * 8;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int x = 0; x < X.N; ++x) {
      const float v = pr[x];
      const float* br = bi + x * X.N + X.lx;
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(v, br[j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) tmp[X.r * 32 + X.c0 + j] = acc[j];
  } else if (active && X.kind == 2) {
    const float* bi = BasisInv(K, 4);
    const float* pr = pix + X.r * 32 + X.c0;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int x = 0; x < 4; ++x) {
      const float v = pr[x];
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(v, bi[x * 4 + j], acc[j]);
    }
    for (int j = 0; j < 4; ++j) tmp[X.r * 32 + X.c0 + j] = acc[j];
  }
  __syncthreads();
  if (active && X.kind == 1) {
    // F[i][j'] with i = ly, j' = lx + j: sum_y Binv[y][j'] U[y][i].
    const float* bi = BasisInv(K, X.N);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int y = 0; y < X.N; ++y) {
      const float u = tmp[(X.oby * 8 + y) * 32 + X.obx * 8 + X.ly];
      const float* br = bi + y * X.N + X.lx;
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(br[j], u, acc[j]);
    }
    for (int j = 0; j < 4; ++j) {
      const int flat = X.ly * X.N + X.lx + j;
      coef[(X.oby + flat / (64 * X.s)) * 256 + X.obx * 64 + flat % (64 * X.s)] =
          acc[j];
    }
  } else if (active && X.kind == 2) {
    const float* bi = BasisInv(K, 4);
    const int sy = X.ly >> 2, sx = X.lx >> 2, i = X.ly & 3;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int y = 0; y < 4; ++y) {
      const float u = tmp[(X.oby * 8 + sy * 4 + y) * 32 + X.obx * 8 + sx * 4 + i];
      for (int j = 0; j < 4; ++j) acc[j] = fmaf(bi[y * 4 + j], u, acc[j]);
    }
    float* cb = coef + X.oby * 256 + X.obx * 64;
    for (int j = 0; j < 4; ++j) cb[(sy + i * 2) * 8 + sx + j * 2] = acc[j];
  } else if (active && X.kind == 3 && X.owner) {
    DirectForward8(X.raw, pix + X.r * 32 + X.c0, 32,
                   coef + X.oby * 256 + X.obx * 64);
  }
  __syncthreads();
  if (active && X.kind == 2 && X.ly == 0 && X.lx == 0) {
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
