// Frame filters (gfx950): min/max of the EPF guide, the edge-preserving
// filter fused with adaptive reconstruction, the dot dictionary, and the
// final Gaborish + XYB->linear conversion.
//
//   MinMax / MakeGuide           pik/epf_target.cc:834-1070
//   SigmaFromQuant               pik/epf_target.cc:1072-1111
//   AdaptiveFilter (scalar form) pik/epf_target.cc:1113-1277, 53-80, 304-376,
//                                669-690
//   AdaptiveReconstruction       pik/adaptive_reconstruction.cc:73-244,
//                                334-471
//   AdaptiveDCReconstruction     pik/adaptive_reconstruction.cc:300-332
//   BlockDictionary::Apply<true> pik/block_dictionary.cc:346-482
//   ConvolveGaborish             pik/gaborish.cc:119-159
//   OpsinToLinear / XybToRgb     pik/opsin_inverse.cc:39-123
#include <float.h>

#include "kernels.h"

namespace pikgpu {

namespace {

__device__ inline float ImgAt(const Img& in, int c, int x, int y) {
  const int mx = min(Mirror(x, in.W), in.vw - 1);
  const int my = min(Mirror(y, in.H), in.vh - 1);
  return in.p[c * in.plane + size_t(my) * in.pitch + mx];
}

// LUTs of SigmaFromQuant (pik/epf_target.cc:1161-1196).
__constant__ float kSigmaLut[32] = {
    1.9815775622811198f, 1.9715084740908622f, 1.6819963065873933f,
    1.2146133632942862f, 1.0395364091521881f, 0.93552327583169714f,
    0.68568655651684773f, 0.51174440217871964f, 0.36397262821018583f,
    0.31621830414136975f, 0.30262954326557712f, 0.246314237855494f,
    0.21617524864418683f, 0.10f, 0.05f, 0.0f,
    float(1.9815775622811198 / 2.0), float(1.9715084740908622 / 2.0),
    float(1.6819963065873933 / 2.0), float(1.2146133632942862 / 2.0),
    float(1.0395364091521881 / 2.0), float(0.93552327583169714 / 2.0),
    float(0.68568655651684773 / 2.0), float(0.51174440217871964 / 2.0),
    float(0.36397262821018583 / 2.0), float(0.31621830414136975 / 2.0),
    float(0.30262954326557712 / 2.0), float(0.246314237855494 / 2.0),
    float(0.21617524864418683 / 2.0), float(0.10 / 2.0), float(0.05 / 2.0),
    float(0.0 / 2.0)};

__device__ inline int SigmaFromQuant(float signal, float stretch, int lut_id) {
  const float* lut = kSigmaLut + 16 * lut_id;
  const float min_signal = 0.022156f;
  const float max_signal = 0.531738f;
  const float mul_signal = 29.435892f;
  float unscaled;
  if (signal <= min_signal) {
    unscaled = lut[0];
  } else if (signal >= max_signal) {
    unscaled = lut[15];
  } else {
    const float pos = (signal - min_signal) * mul_signal;
    const int trunc = int(pos);
    const float frac = pos - float(trunc);
    unscaled = frac * lut[trunc + 1] + (1.0f - frac) * lut[trunc];
  }
  const int sigma = int(unscaled * stretch + 0.5182760822018414f);
  return min(sigma, kMaxSigma);
}

// AcStrategy::ARQuantScale (pik/ac_strategy.h:122-130).
__device__ inline float ARQuantScale(int raw) {
  if (raw == kDCT32) return float(1.2282996852328099);
  if (raw == kDCT16) return float(1.1423171621463439);
  if (raw == kDCT4X4 || raw == kDCT4X4NoHF) return 1.2f;
  if (raw == kDCT2X2) return float(1.0098134203870499);
  return 1.0f;
}

// WeightFast (pik/epf_target.cc:304-376): exponent-field approximation of
// exp(-s * sad) from the per-sigma integer multiplier, for a SAD already
// clamped to kClampedSAD and mul64 = multiplier * 64 (the reference's
// sad << 6 folded into the multiplier; |mul64| < 2^23, |product| < 2^31).
// ((p >> 16) + (127 << 7)) << 16 is evaluated as (p & 0xFFFF0000) +
// (127 << 23), the same bits.
__device__ inline float WeightOfClamped(int sad, int mul64) {
  const int p = __mul24(sad, mul64);
  return __uint_as_float((uint32_t(p) & 0xFFFF0000u) + (127u << 23));
}

typedef unsigned short U16x2 __attribute__((ext_vector_type(2)));

// Lane-wise maximum / minimum of two packed pairs of u16.
__device__ inline uint32_t PkMaxU16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(
      uint32_t, __builtin_elementwise_max(__builtin_bit_cast(U16x2, a),
                                          __builtin_bit_cast(U16x2, b)));
}
__device__ inline uint32_t PkMinU16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(
      uint32_t, __builtin_elementwise_min(__builtin_bit_cast(U16x2, a),
                                          __builtin_bit_cast(U16x2, b)));
}

// Linear sample in [0, 255] -> 8-bit sRGB: the sRGB transfer function
// (IEC 61966-2-1) of the [0, 1]-scaled value, clamped to [0, 1], times 255
// and rounded as the reference's external-image conversion does
// (CastClip01 and ToExternal, pik/external_image.cc:190-197, 666-696).
__device__ inline uint8_t Srgb8(float linear255) {
  const float v = linear255 * (1.0f / 255);
  float e = v <= 0.0031308f ? 12.92f * v
                            : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
  e = fminf(fmaxf(e, 0.0f), 1.0f);
  return static_cast<uint8_t>(static_cast<int>(e * 255.0f + 0.5f));
}

__device__ inline uint32_t LdsDword(const uint8_t* p) {
  return *reinterpret_cast<const uint32_t*>(p);
}

}  // namespace

// ---------------------------------------------------------------------------
// Partial per-channel min/max of the guide over a W x H image (MinMax,
// pik/epf_target.cc:975-1004); one entry of 6 floats per workgroup. Thread
// x walks column x of its row range with the 3x3 window of the Gaborish in
// registers (Symmetric3's operation order, pik/convolve.h:605-621).
__device__ __forceinline__ void KMinMaxPartialBody(
    Img in, int gaborish, FrameParams P, float* partial, const VBlock& vb) {
  __shared__ float s_mm[6][256];
  const int tid = threadIdx.x;
  const int x = vb.x * 256 + tid;
  const int rows = (in.H + vb.ny - 1) / vb.ny;
  const int ya = vb.y * rows, yb = min(in.H, ya + rows);
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  if (x < in.W) {
    // Source columns and rows: mirrored at W x H, then replicated beyond
    // the valid region (ImgAt). The three channels advance together so
    // each row step issues nine independent loads.
    const int xc = min(x, in.vw - 1);
    const int xm = min(Mirror(x - 1, in.W), in.vw - 1);
    const int xp = min(Mirror(x + 1, in.W), in.vw - 1);
    if (!gaborish) {
      for (int y = ya; y < yb; ++y) {
        const size_t r = size_t(min(y, in.vh - 1)) * in.pitch;
        for (int c = 0; c < 3; ++c) {
          const float v = in.p[c * in.plane + r + xc];
          mn[c] = fminf(mn[c], v);
          mx[c] = fmaxf(mx[c], v);
        }
      }
    } else {
      float t[3][3], m[3][3];  // [channel][left, centre, right]
      const size_t r0 = size_t(min(Mirror(ya - 1, in.H), in.vh - 1)) * in.pitch;
      const size_t r1 = size_t(min(ya, in.vh - 1)) * in.pitch;
      for (int c = 0; c < 3; ++c) {
        const float* pc = in.p + c * in.plane;
        t[c][0] = pc[r0 + xm];
        t[c][1] = pc[r0 + xc];
        t[c][2] = pc[r0 + xp];
        m[c][0] = pc[r1 + xm];
        m[c][1] = pc[r1 + xc];
        m[c][2] = pc[r1 + xp];
      }
      // the row below, loaded one iteration ahead
      float nb[3][3];
      {
        const int yn = ya + 1 < in.H ? ya + 1 : in.H - 1;
        const size_t rb = size_t(min(yn, in.vh - 1)) * in.pitch;
        for (int c = 0; c < 3; ++c) {
          const float* pc = in.p + c * in.plane;
          nb[c][0] = pc[rb + xm];
          nb[c][1] = pc[rb + xc];
          nb[c][2] = pc[rb + xp];
        }
      }
      for (int y = ya; y < yb; ++y) {
        float b[3][3];
        for (int c = 0; c < 3; ++c) {
          for (int i = 0; i < 3; ++i) b[c][i] = nb[c][i];
        }
        if (y + 1 < yb) {
          const int yn = y + 2 < in.H ? y + 2 : in.H - 1;
          const size_t rb = size_t(min(yn, in.vh - 1)) * in.pitch;
          for (int c = 0; c < 3; ++c) {
            const float* pc = in.p + c * in.plane;
            nb[c][0] = pc[rb + xm];
            nb[c][1] = pc[rb + xc];
            nb[c][2] = pc[rb + xp];
          }
        }
        for (int c = 0; c < 3; ++c) {
          const float bl = b[c][0], bc = b[c][1], br = b[c][2];
          const float sum_tb = t[c][1] + bc;
          const float mul0 = m[c][1] * P.gab_w[0];
          const float sum_lr = m[c][0] + m[c][2];
          const float x1 = sum_tb + sum_lr;
          const float mul1 = x1 * P.gab_w[1] + mul0;
          const float sum_t2 = t[c][0] + t[c][2];
          const float sum_b2 = bl + br;
          const float x2 = sum_t2 + sum_b2;
          const float v = x2 * P.gab_w[2] + mul1;
          mn[c] = fminf(mn[c], v);
          mx[c] = fmaxf(mx[c], v);
          t[c][0] = m[c][0];
          t[c][1] = m[c][1];
          t[c][2] = m[c][2];
          m[c][0] = bl;
          m[c][1] = bc;
          m[c][2] = br;
        }
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    s_mm[c][tid] = mn[c];
    s_mm[3 + c][tid] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      for (int c = 0; c < 3; ++c) {
        s_mm[c][tid] = fminf(s_mm[c][tid], s_mm[c][tid + s]);
        s_mm[3 + c][tid] = fmaxf(s_mm[3 + c][tid], s_mm[3 + c][tid + s]);
      }
    }
    __syncthreads();
  }
  if (tid < 6) {
    partial[(vb.y * vb.nx + vb.x) * 6 + tid] = s_mm[tid][0];
  }
}
__global__ void __launch_bounds__(256)
    KMinMaxPartial(Img in, int gaborish, FrameParams P, float* partial) {
  KMinMaxPartialBody(in, gaborish, P, partial, PlainBlock());
}

// KMinMaxPartial for an image whose rows are float4-aligned and whose width
// is a multiple of 4 with no replicated margin (vw == W, vh == H): thread t
// walks columns 4t .. 4t + 3 with one float4 and the two mirrored edge
// columns per row and channel, the next row loaded one iteration ahead.
__device__ __forceinline__ void KMinMaxPartial4Body(
    Img in, int gaborish, FrameParams P, float* partial, const VBlock& vb) {
  __shared__ float s_mm[6][256];
  const int tid = threadIdx.x;
  const int x = (vb.x * 256 + tid) * 4;
  const int rows = (in.H + vb.ny - 1) / vb.ny;
  const int ya = vb.y * rows, yb = min(in.H, ya + rows);
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  if (x < in.W && ya < yb) {
    const int xm = Mirror(x - 1, in.W), xp = Mirror(x + 4, in.W);
    // columns x - 1 .. x + 4 of row y, channel c
    auto load_row = [&](int y, int c, float* v) {
      const float* pr = in.p + c * in.plane + size_t(y) * in.pitch;
      const float4 q = *reinterpret_cast<const float4*>(pr + x);
      v[0] = pr[xm];
      v[1] = q.x;
      v[2] = q.y;
      v[3] = q.z;
      v[4] = q.w;
      v[5] = pr[xp];
    };
    if (!gaborish) {
      for (int y = ya; y < yb; ++y) {
        for (int c = 0; c < 3; ++c) {
          const float4 q = *reinterpret_cast<const float4*>(
              in.p + c * in.plane + size_t(y) * in.pitch + x);
          mn[c] = fminf(fminf(mn[c], q.x), fminf(q.y, fminf(q.z, q.w)));
          mx[c] = fmaxf(fmaxf(mx[c], q.x), fmaxf(q.y, fmaxf(q.z, q.w)));
        }
      }
    } else {
      float t[3][6], m[3][6], nb[3][6];
      const int y0 = Mirror(ya - 1, in.H);
      const int y2 = ya + 1 < in.H ? ya + 1 : in.H - 1;
      for (int c = 0; c < 3; ++c) {
        load_row(y0, c, t[c]);
        load_row(ya, c, m[c]);
        load_row(y2, c, nb[c]);
      }
      for (int y = ya; y < yb; ++y) {
        float b[3][6];
        for (int c = 0; c < 3; ++c) {
          for (int i = 0; i < 6; ++i) b[c][i] = nb[c][i];
        }
        if (y + 1 < yb) {
          const int yn = y + 2 < in.H ? y + 2 : in.H - 1;
          for (int c = 0; c < 3; ++c) load_row(yn, c, nb[c]);
        }
        for (int c = 0; c < 3; ++c) {
          for (int k = 0; k < 4; ++k) {
            // strategy::Symmetric3's operation order (pik/convolve.h:605-621)
            const float sum_tb = t[c][k + 1] + b[c][k + 1];
            const float mul0 = m[c][k + 1] * P.gab_w[0];
            const float sum_lr = m[c][k] + m[c][k + 2];
            const float x1 = sum_tb + sum_lr;
            const float mul1 = x1 * P.gab_w[1] + mul0;
            const float sum_t2 = t[c][k] + t[c][k + 2];
            const float sum_b2 = b[c][k] + b[c][k + 2];
            const float x2 = sum_t2 + sum_b2;
            const float v = x2 * P.gab_w[2] + mul1;
            mn[c] = fminf(mn[c], v);
            mx[c] = fmaxf(mx[c], v);
          }
          for (int i = 0; i < 6; ++i) {
            t[c][i] = m[c][i];
            m[c][i] = b[c][i];
          }
        }
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    s_mm[c][tid] = mn[c];
    s_mm[3 + c][tid] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      for (int c = 0; c < 3; ++c) {
        s_mm[c][tid] = fminf(s_mm[c][tid], s_mm[c][tid + s]);
        s_mm[3 + c][tid] = fmaxf(s_mm[3 + c][tid], s_mm[3 + c][tid + s]);
      }
    }
    __syncthreads();
  }
  if (tid < 6) {
    partial[(vb.y * vb.nx + vb.x) * 6 + tid] = s_mm[tid][0];
  }
}
__global__ void __launch_bounds__(256)
    KMinMaxPartial4(Img in, int gaborish, FrameParams P, float* partial) {
  KMinMaxPartial4Body(in, gaborish, P, partial, PlainBlock());
}

// Reduces the partials to mm[0..5] = min[3], max[3]; mm[6] = 255 / range
// (the guide multiplier and the sigma stretch, both 1 for a flat image),
// mm[7] = the minimum over channels.
__device__ __forceinline__ void KMinMaxFinalBody(
    const float* partial, int n, float* mm, const VBlock& vb) {
  __shared__ float s_mm[6][256];
  const int tid = threadIdx.x;
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = tid; i < n; i += 256) {
    for (int c = 0; c < 3; ++c) {
      v[c] = fminf(v[c], partial[i * 6 + c]);
      v[3 + c] = fmaxf(v[3 + c], partial[i * 6 + 3 + c]);
    }
  }
  for (int k = 0; k < 6; ++k) s_mm[k][tid] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      for (int c = 0; c < 3; ++c) {
        s_mm[c][tid] = fminf(s_mm[c][tid], s_mm[c][tid + s]);
        s_mm[3 + c][tid] = fmaxf(s_mm[3 + c][tid], s_mm[3 + c][tid + s]);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    float all_min = s_mm[0][0], all_max = s_mm[3][0];
    for (int c = 0; c < 3; ++c) {
      mm[c] = s_mm[c][0];
      mm[3 + c] = s_mm[3 + c][0];
      all_min = fminf(all_min, s_mm[c][0]);
      all_max = fmaxf(all_max, s_mm[3 + c][0]);
    }
    const float range = all_max - all_min;
    mm[6] = range == 0.0f ? 1.0f : 255.0f / range;
    mm[7] = all_min;
  }
}
__global__ void __launch_bounds__(256)
    KMinMaxFinal(const float* partial, int n, float* mm) {
  KMinMaxFinalBody(partial, n, mm, PlainBlock());
}

// ---------------------------------------------------------------------------
// u8 EPF guide of the IDCT image (the guide AdaptiveFilter builds,
// pik/epf_target.cc:1129-1148): the Gaborish-smoothed image (the image
// itself for the sharpened guide), minus the minimum over channels, times
// the guide stretch, truncated to bytes. One 32x32 tile per workgroup,
// written at (x + 6, y + 6) of the guide image; KGuideRing fills its
// 6-pixel border.
__device__ __forceinline__ void KGuideBody(
    Img in, int gaborish, FrameParams P, FrameBuffers B, const float* mm,
    const VBlock& vb) {
  constexpr int kWP = 35;  // pitch of the 34 x 34 window rows
  __shared__ float s_w[3][34 * kWP];
  const int tid = threadIdx.x;
  const int tiles_x = (in.W + 31) / 32;
  const int tile = XcdTile(vb.x, tiles_x * ((in.H + 31) / 32));
  if (tile < 0) return;
  const int x0 = (tile % tiles_x) * 32, y0 = (tile / tiles_x) * 32;
  // The tile with a 1-pixel mirrored ring (window row / column 0 is pixel
  // -1 of the tile); positions past the ring (a tile wider than the image)
  // are clamped into the image and never used. Tiles whose window columns
  // all lie inside the image load aligned float4s (columns -4 .. +35) and
  // mirror only the rows.
  if (x0 >= 4 && x0 + 36 <= in.W) {
    constexpr int kVec = 3 * 34 * 10;  // float4s per tile
    constexpr int kPer = (kVec + 255) / 256;
    float4 wv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      wv[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (i < kVec) {
        const int row = i / 10, q = i - row * 10;
        const int c = row / 34, ly = row - c * 34;
        int y = y0 - 1 + ly;
        y = y < 0 ? -y - 1 : (y >= in.H ? 2 * in.H - 1 - y : y);
        y = min(max(y, 0), in.H - 1);
        wv[u] = *reinterpret_cast<const float4*>(
            in.p + c * in.plane + size_t(y) * in.pitch + x0 - 4 + 4 * q);
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      if (i < kVec) {
        const int row = i / 10, q = i - row * 10;
        const int c = row / 34, ly = row - c * 34;
        // float4 q holds window columns 4q - 3 .. 4q; the window is
        // columns 0 .. 33
        float* d = s_w[c] + ly * kWP + 4 * q - 3;
        if (q == 0) {
          d[3] = wv[u].w;
        } else if (q == 9) {
          d[0] = wv[u].x;
        } else {
          d[0] = wv[u].x;
          d[1] = wv[u].y;
          d[2] = wv[u].z;
          d[3] = wv[u].w;
        }
      }
    }
  } else {
    constexpr int kWin = 3 * 34 * 34;
    constexpr int kPer = (kWin + 255) / 256;
    float wv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      float v = 0.0f;
      if (i < kWin) {
        const int c = i / (34 * 34), rem = i - c * (34 * 34);
        const int ly = rem / 34, lx = rem - ly * 34;
        int x = x0 - 1 + lx, y = y0 - 1 + ly;
        x = x < 0 ? -x - 1 : (x >= in.W ? 2 * in.W - 1 - x : x);
        y = y < 0 ? -y - 1 : (y >= in.H ? 2 * in.H - 1 - y : y);
        x = min(max(x, 0), in.W - 1);
        y = min(max(y, 0), in.H - 1);
        v = in.p[c * in.plane + size_t(y) * in.pitch + x];
      }
      wv[u] = v;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      if (i < kWin) {
        const int c = i / (34 * 34), rem = i - c * (34 * 34);
        const int ly = rem / 34, lx = rem - ly * 34;
        s_w[c][ly * kWP + lx] = wv[u];
      }
    }
  }
  __syncthreads();
  // Each thread: pixels j0 .. j0 + 3 of tile row i, all three channels.
  const float all_min = mm[7];
  const float gmul = mm[6];
  const size_t gplane = size_t(B.grows) * B.gpitch;
  const int i = tid >> 3, j0 = (tid & 7) << 2;
  const int y = y0 + i, x = x0 + j0;
  if (y >= in.H || x >= in.W) return;
  const int n = min(4, in.W - x);
  // guide byte of pixel (x, y) at (x + 6, y + 6); x + 6 is even
  uint8_t* grow = B.gimg + size_t(y + 6) * B.gpitch + x + 6;
  for (int c = 0; c < 3; ++c) {
    float w[3][6];
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 6; ++k) w[r][k] = s_w[c][(i + r) * kWP + j0 + k];
    }
    uint8_t q[4];
    for (int k = 0; k < 4; ++k) {
      float g;
      if (gaborish) {
        // strategy::Symmetric3's operation order (pik/convolve.h:605-621)
        const float sum_tb = w[0][k + 1] + w[2][k + 1];
        const float mul0 = w[1][k + 1] * P.gab_w[0];
        const float sum_lr = w[1][k] + w[1][k + 2];
        const float x1 = sum_tb + sum_lr;
        const float mul1 = x1 * P.gab_w[1] + mul0;
        const float sum_t2 = w[0][k] + w[0][k + 2];
        const float sum_b2 = w[2][k] + w[2][k + 2];
        const float x2 = sum_t2 + sum_b2;
        g = x2 * P.gab_w[2] + mul1;
      } else {
        g = w[1][k + 1];
      }
      if (B.keep && k < n) {
        B.guide[size_t(c) * P.py * B.ipitch + size_t(y) * B.ipitch + x + k] = g;
      }
      const float scaled = (g - all_min) * gmul;
      q[k] = static_cast<uint8_t>(static_cast<int>(scaled));
    }
    uint8_t* g8 = grow + c * gplane;
    if (n == 4) {
      *reinterpret_cast<uint16_t*>(g8) = uint16_t(q[0] | (q[1] << 8));
      *reinterpret_cast<uint16_t*>(g8 + 2) = uint16_t(q[2] | (q[3] << 8));
    } else {
      for (int k = 0; k < n; ++k) g8[k] = q[k];
    }
  }
}
__global__ void __launch_bounds__(256)
    KGuide(Img in, int gaborish, FrameParams P, FrameBuffers B,
           const float* mm) {
  KGuideBody(in, gaborish, P, B, mm, PlainBlock());
}

// The 6-pixel border of the guide image: each position takes the byte of
// the pixel it mirrors. (The guide of a mirrored position equals its
// source pixel's bit for bit: the 3x3 sums of the mirrored window only
// exchange the operands of commutative additions.)
__device__ __forceinline__ void KGuideRingBody(
    FrameParams P, FrameBuffers B, const VBlock& vb) {
  const int W = P.px, H = P.py;
  const int pw = W + 12;
  const int n_tb = 6 * pw, n_lr = 6 * H;
  const int per_c = 2 * n_tb + 2 * n_lr;
  const int k = vb.x * 256 + threadIdx.x;
  if (k >= 3 * per_c) return;
  const int c = k / per_c;
  int r = k - c * per_c;
  int X, Y;
  if (r < n_tb) {
    Y = r / pw;
    X = r - Y * pw;
  } else if ((r -= n_tb) < n_tb) {
    Y = H + 6 + r / pw;
    X = r - (r / pw) * pw;
  } else if ((r -= n_tb) < n_lr) {
    Y = 6 + r / 6;
    X = r - (r / 6) * 6;
  } else {
    r -= n_lr;
    Y = 6 + r / 6;
    X = W + 6 + (r - (r / 6) * 6);
  }
  const int sx = Mirror(X - 6, W), sy = Mirror(Y - 6, H);
  uint8_t* g = B.gimg + size_t(c) * B.grows * B.gpitch;
  g[size_t(Y) * B.gpitch + X] = g[size_t(sy + 6) * B.gpitch + sx + 6];
}
__global__ void __launch_bounds__(256)
    KGuideRing(FrameParams P, FrameBuffers B) {
  KGuideRingBody(P, B, PlainBlock());
}

// ---------------------------------------------------------------------------
// EPF (+ AR clamp) on one 32x32 tile. dc_mode: the image is the 8-padded DC
// image (constant quant field QuantDC/2, lut 0, DCT8 everywhere), no AR, and
// the result is clamped to +-half a DC step and written to B.dc.
//
// LDS: the input window (46 x 46 per channel) is dead once the filter has
// run -- each thread keeps the 12 original values it needs for AR in
// registers -- so AR's coefficients and transform scratch reuse it; with
// ~45 KB per workgroup three workgroups share a CU.
//
// Image mode (!kDc) reads the u8 guide from the guide image KGuide wrote;
// the DC mode quantises its input itself (no Gaborish on the DC image).
template <bool kDc>
__device__ __forceinline__ void KEpfBody(
    Img in, FrameParams P, FrameBuffers B, const DeviceConsts* K,
    const float* mm, float* out, int out_pitch, size_t out_plane,
    const VBlock& vb) {
  constexpr int dc_mode = kDc;
  // Input window: the DC mode keeps the tile with a 7-pixel border (its
  // guide is computed from it); the image mode only the filter's window,
  // pixels -3 .. +35 in x and -3 .. +34 in y. Window row / column 0 is
  // pixel -kOrg of the tile.
  constexpr int kOrg = kDc ? 7 : 3;
  constexpr int kRows = kDc ? 46 : 38;
  constexpr int kCols = kDc ? 46 : 39;
  constexpr int kInP = kDc ? 48 : 40;  // pitch of s_in rows
  constexpr int kGP = 48;              // pitch of s_g rows (44 used)
  constexpr int kInPlane = kRows * kInP;
  constexpr int kGuideFloats = 3 * 44 * kGP / 4;
  constexpr int kBig = 3 * kInPlane + kGuideFloats > 3 * (kCoefPlane + kTilePlane)
                           ? 3 * kInPlane + kGuideFloats
                           : 3 * (kCoefPlane + kTilePlane);
  // s_in and s_g, then (AR) s_coef and s_tmp; ~39 KB in the image mode:
  // four workgroups per CU.
  __shared__ __attribute__((aligned(16))) float s_big[kBig];
  __shared__ float s_out[3][kTilePlane];
  __shared__ uint8_t s_acs[16];
  __shared__ int s_sigma[16];
  __shared__ TileTasks s_tasks;
  float (*s_in)[kInPlane] = reinterpret_cast<float (*)[kInPlane]>(s_big);
  uint8_t (*s_g)[44 * kGP] =
      reinterpret_cast<uint8_t (*)[44 * kGP]>(s_big + 3 * kInPlane);

  const int tid = threadIdx.x;
  const int tiles_x = (in.W + 31) / 32;
  const int tile = XcdTile(vb.x, tiles_x * ((in.H + 31) / 32));
  if (tile < 0) return;
  const int x0 = (tile % tiles_x) * 32, y0 = (tile / tiles_x) * 32;
  const float all_min = mm[7];
  const float gmul = mm[6];

  // AR (step 5) needs the descriptor of the storage block whose
  // coefficients this thread owns (the matrix slots and 1 / qf of the
  // transform's first block); it loads with the input window.
  const int crow = tid >> 6, ccol = (tid & 63) << 2;
  const int ar_blk =
      (y0 / 8 + crow) * P.bx + x0 / 8 + (ccol >> 6);
  const bool ar_valid =
      !dc_mode && x0 / 8 + (ccol >> 6) < P.bx && y0 / 8 + crow < P.by;
  uint4 ar_desc = make_uint4(0, 0, 0, 0);
  if (ar_valid) ar_desc = B.desc[2 * ar_blk + 1];
  // inputs of the per-block sigma (step 2), loaded with the window
  uint8_t sg_acs = 0;
  int sg_qf = 0, sg_lut = 0;
  const bool sg_valid = tid < 16 && (x0 / 8 + (tid & 3)) * 8 < in.W &&
                        (y0 / 8 + (tid >> 2)) * 8 < in.H;
  if (!kDc && sg_valid) {
    const int blk = (y0 / 8 + (tid >> 2)) * P.bx + x0 / 8 + (tid & 3);
    sg_acs = B.acs[blk];
    sg_qf = B.qf[blk];
    sg_lut = B.lut[blk];
  }

  // 1. input tile with a 7-pixel mirrored border (the DC mode computes its
  //    guide from it); the image mode only needs the filter's window,
  //    pixels -3 .. +35 of the tile in x and -3 .. +34 in y. One reflection
  //    gives the mirror for every position within 7 of the image (both
  //    images are at least 8 wide and high); positions further out (a tile
  //    wider than a small image reaches up to 38 past its origin) are only
  //    clamped into the image, their values are never used. All loads are
  //    issued before the LDS stores. Image-mode tiles whose window columns
  //    all lie inside the image load aligned float4s (columns -4 .. +35,
  //    the first one dropped) and mirror only the rows.
  if (!kDc && x0 >= 4 && x0 + 36 <= in.W && in.vw == in.W) {
    constexpr int kVec = 3 * kRows * 10;  // float4s per tile
    constexpr int kPer = (kVec + 255) / 256;
    float4 wv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      wv[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (i < kVec) {
        const int row = i / 10, q = i - row * 10;
        const int c = row / kRows, ly = row - c * kRows;
        int y = y0 - kOrg + ly;
        y = y < 0 ? -y - 1 : (y >= in.H ? 2 * in.H - 1 - y : y);
        y = min(max(y, 0), in.vh - 1);
        wv[u] = *reinterpret_cast<const float4*>(
            in.p + c * in.plane + size_t(y) * in.pitch + x0 - 4 + 4 * q);
      }
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      if (i < kVec) {
        const int row = i / 10, q = i - row * 10;
        const int c = row / kRows, ly = row - c * kRows;
        float* d = s_in[c] + ly * kInP + 4 * q - 1;
        if (q > 0) d[0] = wv[u].x;
        d[1] = wv[u].y;
        d[2] = wv[u].z;
        d[3] = wv[u].w;
      }
    }
  } else {
    constexpr int kWin = 3 * kRows * kCols;
    constexpr int kPer = (kWin + 255) / 256;
    float wv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      float v = 0.0f;
      if (i < kWin) {
        const int c = i / (kRows * kCols), rem = i - c * (kRows * kCols);
        const int ly = rem / kCols, lx = rem - ly * kCols;
        int x = x0 - kOrg + lx, y = y0 - kOrg + ly;
        x = x < 0 ? -x - 1 : (x >= in.W ? 2 * in.W - 1 - x : x);
        y = y < 0 ? -y - 1 : (y >= in.H ? 2 * in.H - 1 - y : y);
        x = min(max(x, 0), in.vw - 1);
        y = min(max(y, 0), in.vh - 1);
        v = in.p[c * in.plane + size_t(y) * in.pitch + x];
      }
      wv[u] = v;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * 256;
      if (i < kWin) {
        const int c = i / (kRows * kCols), rem = i - c * (kRows * kCols);
        const int ly = rem / kCols, lx = rem - ly * kCols;
        s_in[c][ly * kInP + lx] = wv[u];
      }
    }
  }
  if (!kDc) {
    // the guide window: 44 rows of 11 aligned dwords per channel (the guide
    // image's 6-pixel border makes window column 0 the padded column x0)
    const size_t gplane = size_t(B.grows) * B.gpitch;
    for (int k = tid; k < 3 * 44 * 11; k += 256) {
      const int c = k / (44 * 11), rem = k - c * (44 * 11);
      const int row = rem / 11, d = rem - row * 11;
      const uint32_t v = *reinterpret_cast<const uint32_t*>(
          B.gimg + c * gplane + size_t(y0 + row) * B.gpitch + x0 + 4 * d);
      *reinterpret_cast<uint32_t*>(&s_g[c][row * kGP + 4 * d]) = v;
    }
  }
  // 2. per-block sigma and strategy
  if (tid < 16) {
    int sigma = 0;
    uint8_t a = 0xFF;
    if (sg_valid) {
      float ac_q, scale;
      int lut_id;
      if (dc_mode) {
        ac_q = float(P.dc_qf);
        scale = 1.0f;
        lut_id = 0;
        a = 0;
      } else {
        a = sg_acs;
        ac_q = float(sg_qf);
        scale = ARQuantScale(a >> 4);
        lut_id = sg_lut;
      }
      const float quant = ac_q * scale * 0.039324273f;
      sigma = SigmaFromQuant(quant, gmul, lut_id);
    }
    s_sigma[tid] = sigma;
    s_acs[tid] = a;
  }
  __syncthreads();
  // 3. DC mode: u8 guide over the tile plus a 6-pixel border (MakeGuide,
  //    no Gaborish on the DC image).
  if (kDc) {
    for (int i = tid; i < 3 * 44 * 44; i += 256) {
      const int c = i / (44 * 44), rem = i % (44 * 44);
      const int ly = rem / 44, lx = rem % 44;
      const float g = s_in[c][(ly + 1) * kInP + (lx + 1)];
      const float scaled = (g - all_min) * gmul;
      s_g[c][ly * kGP + lx] = static_cast<uint8_t>(static_cast<int>(scaled));
    }
    __syncthreads();
  }

  // 4. EPF (AdaptiveFilter, pik/epf_target.cc:1113-1277): each thread
  //    filters 4 horizontally adjacent pixels of one block row. For every
  //    neighbour row offset cy the SADs of the 3x4 patches against the 8
  //    neighbours cx = -3..4 come from two quad-SAD instructions per patch
  //    row and channel (v_qsad_pk_u16_u8: four byte-shifted 4-byte windows
  //    of an 8-byte source against the 4-byte reference patch row, 16-bit
  //    accumulators); channels are combined by their maximum (packed u16),
  //    then WeightFast, the flush below 0.05 and the weighted mean of the
  //    input follow in the reference's sequential neighbour order.
  const int r = tid >> 3, c0 = (tid & 7) << 2;
  {
    const int sigma = s_sigma[(r >> 3) * 4 + (c0 >> 3)];
    const bool inside = x0 + c0 < in.W && y0 + r < in.H;
    if (inside && sigma < kMinSigma) {
      for (int c = 0; c < 3; ++c) {
        for (int j = 0; j < 4; ++j) {
          s_out[c][PixAt(r, c0 + j)] =
              s_in[c][(r + kOrg) * kInP + c0 + kOrg + j];
        }
      }
    } else if (inside) {
      // mul * 64 as a 24-bit signed value (the table holds -32512 .. -774),
      // so the weight multiply is a 24-bit one
      const int mul64 = (K->mul_table[sigma] * 64 << 8) >> 8;
      // The three patch rows iy = 0..2 of a pixel in row r are guide rows
      // r+5+iy (the guide tile has a 6-pixel border); the SAD of a patch is
      // the sum of its rows' SADs. Rows r = 2k and 2k+1 sit in lanes L and
      // L+8 of one 16-lane DPP row and share patch rows 2k+5+1 and 2k+5+2:
      // each lane sums its own row (iy = 1) with its partner's own row
      // (a DPP rotate by 8 within the row, added as one 32-bit word per
      // pair of u16 SADs: a three-row SAD is at most 3060, so no carry
      // crosses the halves), then adds its outer row (iy = 0 for even r,
      // iy = 2 for odd r). Both lanes of a pair are always active together:
      // they share the block (sigma), the columns and, the image height
      // being a multiple of 8, the row test.
      const int oiy = (r & 1) ? 2 : 0;
      // Reference patch rows of pixel j: guide bytes c0+5+j .. c0+8+j of
      // the own row (iy = 1) and the outer row (iy = oiy).
      uint32_t ref_own[3][4], ref_out[3][4];
      for (int ch = 0; ch < 3; ++ch) {
        for (int k = 0; k < 2; ++k) {
          const int iy = k == 0 ? 1 : oiy;
          const uint8_t* row = s_g[ch] + (r + 5 + iy) * kGP + c0 + 4;
          const uint32_t d0 = LdsDword(row), d1 = LdsDword(row + 4);
          for (int j = 0; j < 4; ++j) {
            const uint32_t v =
                j == 3 ? d1 : __builtin_amdgcn_alignbyte(d1, d0, j + 1);
            if (k == 0) ref_own[ch][j] = v;
            else ref_out[ch][j] = v;
          }
        }
      }
      float ws[4][3], sw[4];
      for (int j = 0; j < 4; ++j) {
        ws[j][0] = ws[j][1] = ws[j][2] = 0.0f;
        sw[j] = 0.0f;
      }
      for (int cy = -3; cy <= 3; ++cy) {
        // msad[j][w]: packed u16 SADs of pixel j, words w = cx pairs
        // (-3,-2), (-1,0), (1,2), (3,4).
        uint32_t msad[4][4];
        for (int ch = 0; ch < 3; ++ch) {
          uint64_t acc[4][2];
          for (int k = 0; k < 2; ++k) {
            // k = 0: own patch row, then the pair's sum; k = 1: outer row
            const int iy = k == 0 ? 1 : oiy;
            const uint8_t* row = s_g[ch] + (r + 5 + cy + iy) * kGP + c0;
            const uint32_t d[5] = {LdsDword(row), LdsDword(row + 4),
                                   LdsDword(row + 8), LdsDword(row + 12),
                                   LdsDword(row + 16)};
            for (int j = 0; j < 4; ++j) {
              // windows of neighbours cx = -3..0 start at byte c0+2+j,
              // those of cx = 1..4 at byte c0+6+j
              const int o = 2 + j;
              // (a zero byte shift is the dword itself)
              const int sh = o & 3, q0 = o >> 2;
              const uint32_t a0 =
                  sh == 0 ? d[q0] : __builtin_amdgcn_alignbyte(d[q0 + 1], d[q0], sh);
              const uint32_t a1 = sh == 0 ? d[q0 + 1]
                                          : __builtin_amdgcn_alignbyte(
                                                d[q0 + 2], d[q0 + 1], sh);
              const uint32_t a2 = sh == 0 ? d[q0 + 2]
                                          : __builtin_amdgcn_alignbyte(
                                                d[q0 + 3], d[q0 + 2], sh);
              const uint64_t lo = uint64_t(a0) | (uint64_t(a1) << 32);
              const uint64_t hi = uint64_t(a1) | (uint64_t(a2) << 32);
              if (k == 0) {
                acc[j][0] = __builtin_amdgcn_qsad_pk_u16_u8(lo, ref_own[ch][j], 0);
                acc[j][1] = __builtin_amdgcn_qsad_pk_u16_u8(hi, ref_own[ch][j], 0);
                for (int h = 0; h < 2; ++h) {
                  uint32_t w[2] = {uint32_t(acc[j][h]), uint32_t(acc[j][h] >> 32)};
                  for (int e = 0; e < 2; ++e) {
                    // + the partner lane's word (row_ror:8)
                    w[e] += uint32_t(__builtin_amdgcn_update_dpp(
                        0, int(w[e]), 0x128, 0xF, 0xF, false));
                  }
                  acc[j][h] = uint64_t(w[0]) | (uint64_t(w[1]) << 32);
                }
              } else {
                acc[j][0] = __builtin_amdgcn_qsad_pk_u16_u8(lo, ref_out[ch][j],
                                                            acc[j][0]);
                acc[j][1] = __builtin_amdgcn_qsad_pk_u16_u8(hi, ref_out[ch][j],
                                                            acc[j][1]);
              }
            }
          }
          for (int j = 0; j < 4; ++j) {
            const uint32_t w4[4] = {uint32_t(acc[j][0]),
                                    uint32_t(acc[j][0] >> 32),
                                    uint32_t(acc[j][1]),
                                    uint32_t(acc[j][1] >> 32)};
            for (int w = 0; w < 4; ++w) {
              msad[j][w] = ch == 0 ? w4[w] : PkMaxU16(msad[j][w], w4[w]);
            }
          }
        }
        // the input row of this cy, columns c0-3 .. c0+7 (+1 pad)
        float v[3][12];
        for (int c = 0; c < 3; ++c) {
          const float4* src = reinterpret_cast<const float4*>(
              s_in[c] + (r + kOrg + cy) * kInP + c0 + kOrg - 3);
          for (int q = 0; q < 3; ++q) {
            const float4 t = src[q];
            v[c][4 * q] = t.x;
            v[c][4 * q + 1] = t.y;
            v[c][4 * q + 2] = t.z;
            v[c][4 * q + 3] = t.w;
          }
        }
        for (int cx = -3; cx <= 4; ++cx) {
          const int w = (cx + 3) >> 1, hi = (cx + 3) & 1;
          for (int j = 0; j < 4; ++j) {
            const uint32_t word = PkMinU16(msad[j][w], 0x01FB01FBu);
            const uint32_t sad = hi ? (word >> 16) : (word & 0xFFFFu);
            float wt = WeightOfClamped(int(sad), mul64);
            wt = wt >= 0.05f ? wt : 0.0f;
            for (int c = 0; c < 3; ++c) {
              ws[j][c] = ws[j][c] + v[c][j + cx + 3] * wt;
            }
            sw[j] = sw[j] + wt;
          }
        }
      }
      for (int j = 0; j < 4; ++j) {
        for (int c = 0; c < 3; ++c) {
          s_out[c][PixAt(r, c0 + j)] = ws[j][c] / sw[j];
        }
      }
    }
  }
  __syncthreads();

  if (dc_mode) {
    // Clamp to +-half a quantization step around the dequantised DC
    // (pik/adaptive_reconstruction.cc:320-331); only the unpadded region.
    const int y = y0 + r;
    for (int j = 0; j < 4; ++j) {
      const int x = x0 + c0 + j;
      if (x < P.bx && y < P.by) {
        for (int c = 0; c < 3; ++c) {
          const float dc = s_in[c][(r + kOrg) * kInP + c0 + kOrg + j];
          const float hs = P.dc_half_step[c];
          const float lo = dc - hs, hi = dc + hs;
          const float d = s_out[c][PixAt(r, c0 + j)];
          const float mn = hi < d ? hi : d;
          out[c * out_plane + size_t(y) * out_pitch + x] = lo < mn ? mn : lo;
        }
      }
    }
    return;
  }
  if (B.keep) {
    const int y = y0 + r;
    for (int j = 0; j < 4; ++j) {
      const int x = x0 + c0 + j;
      if (x < in.W && y < in.H) {
        for (int c = 0; c < 3; ++c) {
          B.epf[size_t(c) * P.py * B.ipitch + size_t(y) * B.ipitch + x] =
              s_out[c][PixAt(r, c0 + j)];
        }
      }
    }
  }

  // 5. AR: clamp the DCT of (filtered - original) to half the quantization
  //    interval with one ratio per coefficient shared by X, Y, B. The
  //    matrix entries of this thread's coefficients load during the forward
  //    transform.
  float ar_m[3][4];
  if (ar_valid) {
    const uint32_t slot[3] = {ar_desc.x, ar_desc.y, ar_desc.z};
    for (int c = 0; c < 3; ++c) {
      for (int j = 0; j < 4; ++j) ar_m[c][j] = B.table[slot[c] + (ccol & 63) + j];
    }
  }
  float orig[3][4];
  for (int c = 0; c < 3; ++c) {
    for (int j = 0; j < 4; ++j) {
      orig[c][j] = s_in[c][(r + kOrg) * kInP + c0 + kOrg + j];
      s_out[c][PixAt(r, c0 + j)] = s_out[c][PixAt(r, c0 + j)] - orig[c][j];
    }
  }
  // (BuildTasks ends with a barrier: every thread has read s_in, so the
  // coefficient and scratch buffers may overwrite it.)
  BuildTasks(s_acs, false, &s_tasks);
  const ThreadTasks my_tasks = AssignTasks(s_tasks);
  float (*s_coef)[kCoefPlane] = reinterpret_cast<float (*)[kCoefPlane]>(s_big);
  float (*s_tmp)[kTilePlane] =
      reinterpret_cast<float (*)[kTilePlane]>(s_big + 3 * kCoefPlane);
  TileForward3(my_tasks, &s_out[0][0], &s_coef[0][0], &s_tmp[0][0]);
  {
    // coefficient entries owned by this thread; the same thread scales all
    // three channels of its entries. The quantisation interval uses the
    // quant value of the transform's first block (ar_desc: its
    // inv_global_scale / qf and its matrix slots for this block, so entry
    // k of this storage block is the matrix's flat index idx * 64 + k).
    const uint8_t ca = s_acs[crow * 4 + (ccol >> 6)];
    if (ca != 0xFF) {
      const int raw = ca >> 4, idx = ca & 15;
      const int s = AcsCovered(raw);
      const int N = 8 * s;
      const int lgs = s >> 1;  // log2(s): s is 1, 2 or 4
      const int coby = crow - (idx >> lgs), cobx = (ccol >> 6) - (idx & (s - 1));
      const int cflat0 = (crow - coby) * 64 * s + (ccol - cobx * 64);
      const float inv_quant_ac = __uint_as_float(ar_desc.w);
      for (int j = 0; j < 4; ++j) {
        const int flat = cflat0 + j;
        const int fi = flat >> (3 + lgs), fj = flat & (N - 1);
        float ratio = 1.0f;
        for (int c = 0; c < 3; ++c) {
          float interval;
          if (fi < s && fj < s) {
            // P.mul_dc[c] = table[offsets[c]] * inv_quant_dc
            const float lscale = s == 1 ? 1.0f : K->ascale[s == 2 ? 3 : 4][fi][fj];
            interval = lscale * P.mul_dc[c];
          } else {
            interval = ar_m[c][j] * inv_quant_ac;
          }
          const float upper = 0.5f * interval;
          const float lower = -0.5f * interval;
          const float corr = s_coef[c][CoefAt(crow, ccol + j)];
          const float m1 = lower < corr ? corr : lower;
          const float clamped = m1 < upper ? m1 : upper;
          // An entry inside the interval (clamped == corr != 0) has the
          // ratio corr / corr = 1 exactly, which leaves ratio (<= 1) as it
          // is: only clamped and zero entries need the division.
          if (!(clamped == corr && corr != 0.0f)) {
            const float divisor = __float_as_uint(corr) == 0u ? 1E-7f : corr;
            const float cr = clamped / divisor;
            ratio = cr < ratio ? cr : ratio;
          }
        }
        for (int c = 0; c < 3; ++c) {
          s_coef[c][CoefAt(crow, ccol + j)] =
              s_coef[c][CoefAt(crow, ccol + j)] * ratio;
        }
      }
    }
  }
  __syncthreads();
  TileInverse3(my_tasks, &s_coef[0][0], &s_out[0][0], &s_tmp[0][0]);
  const int y = y0 + r;
  if (y < in.H) {
    for (int c = 0; c < 3; ++c) {
      float* dst = out + c * out_plane + size_t(y) * out_pitch;
      for (int j = 0; j < 4; ++j) {
        const int x = x0 + c0 + j;
        if (x < in.W) dst[x] = s_out[c][PixAt(r, c0 + j)] + orig[c][j];
      }
    }
  }
}
template <bool kDc>
__global__ void __launch_bounds__(256, kDc ? 3 : 4)
    KEpf(Img in, FrameParams P, FrameBuffers B, const DeviceConsts* K,
         const float* mm, float* out, int out_pitch, size_t out_plane) {
  KEpfBody<kDc>(in, P, B, K, mm, out, out_pitch, out_plane, PlainBlock());
}

// ---------------------------------------------------------------------------
// Dot patches (BlockDictionary::Apply<true>, pik/block_dictionary.cc:
// 346-482) on the AR output, in place. Workgroup t owns the 32x32 tile
// dot_tile_ids[t]; each thread owns four pixels of it and adds the tile's
// dots (listed in the reference's application order) to them one dot after
// the other, so every pixel receives its additions in the reference's order.
__device__ __forceinline__ void KDotsBody(
    FrameParams P, FrameBuffers B, const VBlock& vb) {
  __shared__ float s_blk[3][21 * 21];
  const int tid = threadIdx.x;
  const int tile = B.dot_tile_ids[vb.x];
  const int tx0 = (tile % P.dot_tiles_x) * 32, ty0 = (tile / P.dot_tiles_x) * 32;
  const int px = tx0 + (tid & 31);
  const int begin = B.dot_tile_start[vb.x];
  const int end = B.dot_tile_start[vb.x + 1];
  const float kBlockInvScale = 1.0f / 3.5f;
  const float kXybRadius[3] = {0.0301006f, 0.4512295f, 0.4716444f};
  const size_t plane = size_t(P.py) * B.ipitch;
  for (int li = begin; li < end; ++li) {
    const pikgpu_dot& d = B.dots[B.dot_list[li]];  // read in place
    const int xs = d.xsize, ys = d.ysize;
    if (!d.transform) {
      for (int k = 0; k < 4; ++k) {
        const int py = ty0 + (tid >> 5) + 8 * k;
        const int ix = px - int(d.x), iy = py - int(d.y);
        if (ix < 0 || iy < 0 || ix >= xs || iy >= ys || px >= P.px ||
            py >= P.py) {
          continue;
        }
        for (int c = 0; c < 3; ++c) {
          const float val =
              kBlockInvScale * kXybRadius[c] * float(d.pixels[c][iy * xs + ix]);
          float* p = B.ar + c * plane + size_t(py) * B.ipitch + px;
          *p = *p + val;
        }
      }
      continue;
    }
    // 2x nearest-neighbour upsampling with a 2x2 box
    // (pik/block_dictionary.cc:382-401), then a bilinear splat through the
    // inverse affine map.
    const int bw = 2 * xs + 5;
    __syncthreads();
    for (int i = tid; i < 3 * 21 * 21; i += 256) (&s_blk[0][0])[i] = 0.0f;
    __syncthreads();
    for (int i = tid; i < 3 * (2 * ys + 1) * (2 * xs + 1); i += 256) {
      const int c = i / ((2 * ys + 1) * (2 * xs + 1));
      const int rem = i % ((2 * ys + 1) * (2 * xs + 1));
      const int iy = rem / (2 * xs + 1), ix = rem % (2 * xs + 1);
      float val = 0.0f;
      for (int dy = -1; dy < 1; dy++) {
        for (int dx = -1; dx < 1; dx++) {
          int sy = (iy + dy) / 2;
          if (sy < 0) sy = 0;
          if (sy >= ys) sy = ys - 1;
          int sx = (ix + dx) / 2;
          if (sx < 0) sx = 0;
          if (sx >= xs) sx = xs - 1;
          val += float(d.pixels[c][sy * xs + sx]);
        }
      }
      s_blk[c][(iy + 2) * bw + (ix + 2)] = val * 0.25f;
    }
    __syncthreads();
    // BlockPosition keeps dx, dy, width as int64 and the patch size as
    // size_t; the expressions below keep those conversions.
    const float by00 = float(d.y);
    const float bx00 = float(d.x);
    const float bx01 = bx00 + float(d.dx);
    const float by01 = by00 + float(d.dy);
    const float xnorm =
        float(sqrt(double(int64_t(d.dx) * d.dx + int64_t(d.dy) * d.dy)));
    const float ynorm = float(d.width) * 0.5f;
    const float invxnorm = 1.0f / xnorm;
    const float invynorm = 1.0f / ynorm;
    const float deltax = float(d.dy) * invxnorm * ynorm;
    const float deltay = float(-d.dx) * invxnorm * ynorm;
    const float bx11 = bx01 + deltax;
    const float by11 = by01 + deltay;
    const float bx10 = bx00 + deltax;
    const float by10 = by00 + deltay;
    const float inv_det = 1.0f / (float(d.dx) * deltay - float(d.dy) * deltax);
    const float it0 = float(xs) * deltay * inv_det;
    const float it1 = -float(xs) * deltax * inv_det;
    const float it2 = -float(ys) * float(d.dy) * inv_det;
    const float it3 = float(uint64_t(ys) * uint64_t(int64_t(d.dx))) * inv_det;
    int64_t min_x = int64_t(fminf(fminf(bx00, bx01), fminf(bx10, bx11)));
    int64_t min_y = int64_t(fminf(fminf(by00, by01), fminf(by10, by11)));
    int64_t max_x = int64_t(fmaxf(fmaxf(bx00, bx01), fmaxf(bx10, bx11)) + 1);
    int64_t max_y = int64_t(fmaxf(fmaxf(by00, by01), fmaxf(by10, by11)) + 1);
    if (min_x < 0) min_x = 0;
    if (min_y < 0) min_y = 0;
    if (max_x > P.px) max_x = P.px;
    if (max_y > P.py) max_y = P.py;
    const float margin_x = 0.3f * invxnorm;
    const float margin_y = 0.3f * invynorm;
    for (int k = 0; k < 4; ++k) {
      const int py = ty0 + (tid >> 5) + 8 * k;
      if (px < min_x || px >= max_x || py < min_y || py >= max_y) continue;
      const float x = float(px) - bx00 + 0.5f;
      const float y = float(py) - by00 + 0.5f;
      float ox = it0 * x + it1 * y;
      float oy = it2 * x + it3 * y;
      if (ox >= -margin_x && ox < float(xs) + margin_x && oy >= -margin_y &&
          oy < float(ys) + margin_y) {
        ox = 2 * ox + 2;
        oy = 2 * oy + 2;
        const int floorx = int(ox);
        const int ceilx = floorx + 1;
        const float fracx = ox - float(floorx);
        const int floory = int(oy);
        const int ceily = floory + 1;
        const float fracy = oy - float(floory);
        for (int c = 0; c < 3; ++c) {
          const float* blk = s_blk[c];
          const float val =
              (fracx * fracy * blk[ceily * bw + ceilx] +
               fracx * (1.f - fracy) * blk[floory * bw + ceilx] +
               (1.f - fracx) * fracy * blk[ceily * bw + floorx] +
               (1.f - fracx) * (1.f - fracy) * blk[floory * bw + floorx]) *
              kBlockInvScale * kXybRadius[c];
          float* p = B.ar + c * plane + size_t(py) * B.ipitch + px;
          *p = *p + val;
        }
      }
    }
  }
}
__global__ void __launch_bounds__(256)
    KDots(FrameParams P, FrameBuffers B) {
  KDotsBody(P, B, PlainBlock());
}

// ---------------------------------------------------------------------------
// Final Gaborish (on the 8-padded image) and XYB -> linear RGB of the
// xsize x ysize output (pik/compressed_image.cc:1164-1179). Thread x of a
// workgroup walks kFinalRows rows of column x of its 256-column strip with
// the 3x3 window of each channel in registers, so every input value is
// loaded three times (left, centre, right) instead of nine.
__device__ __forceinline__ void KFinalBody(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, const VBlock& vb) {
  const int strips = (P.px + 255) / 256;
  const int tile =
      XcdTile(vb.x, strips * ((P.py + kFinalRows - 1) / kFinalRows));
  if (tile < 0) return;
  const int x = (tile % strips) * 256 + threadIdx.x;
  const int ya = (tile / strips) * kFinalRows;
  const int yb = min(P.py, ya + kFinalRows);
  if (x >= P.px) return;
  const size_t plane = size_t(P.py) * B.ipitch;
  const int xm = Mirror(x - 1, P.px), xp = Mirror(x + 1, P.px);
  // in registers: the output stores could alias K as far as the compiler
  // knows, which would reload the matrix every row
  float inv[9];
  for (int i = 0; i < 9; ++i) inv[i] = K->opsin_inv[i];
  // rows above (t) and at (m) the current row: [channel][left, centre, right]
  float t[3][3], m[3][3];
  {
    const size_t rt = size_t(Mirror(ya - 1, P.py)) * B.ipitch;
    const size_t rm = size_t(ya) * B.ipitch;
    for (int c = 0; c < 3; ++c) {
      const float* src = B.ar + c * plane;
      t[c][0] = src[rt + xm];
      t[c][1] = src[rt + x];
      t[c][2] = src[rt + xp];
      m[c][0] = src[rm + xm];
      m[c][1] = src[rm + x];
      m[c][2] = src[rm + xp];
    }
  }
  // The row below is loaded one iteration ahead, before this row's stores
  // (which the compiler must assume may alias the input).
  float nb[3][3];
  {
    const size_t rb = size_t(ya + 1 < P.py ? ya + 1 : P.py - 1) * B.ipitch;
    for (int c = 0; c < 3; ++c) {
      const float* src = B.ar + c * plane;
      nb[c][0] = src[rb + xm];
      nb[c][1] = src[rb + x];
      nb[c][2] = src[rb + xp];
    }
  }
  for (int y = ya; y < yb; ++y) {
    float b[3][3];
    for (int c = 0; c < 3; ++c) {
      for (int i = 0; i < 3; ++i) b[c][i] = nb[c][i];
    }
    if (y + 1 < yb) {
      const size_t rb = size_t(y + 2 < P.py ? y + 2 : P.py - 1) * B.ipitch;
      for (int c = 0; c < 3; ++c) {
        const float* src = B.ar + c * plane;
        nb[c][0] = src[rb + xm];
        nb[c][1] = src[rb + x];
        nb[c][2] = src[rb + xp];
      }
    }
    float v[3];
    for (int c = 0; c < 3; ++c) {
      const float bl = b[c][0], bc = b[c][1], br = b[c][2];
      if (P.gaborish) {
        // strategy::Symmetric3's operation order (pik/convolve.h:605-621)
        const float sum_tb = t[c][1] + bc;
        const float mul0 = m[c][1] * P.gab_w[0];
        const float sum_lr = m[c][0] + m[c][2];
        const float x1 = sum_tb + sum_lr;
        const float mul1 = x1 * P.gab_w[1] + mul0;
        const float sum_t2 = t[c][0] + t[c][2];
        const float sum_b2 = bl + br;
        const float x2 = sum_t2 + sum_b2;
        v[c] = x2 * P.gab_w[2] + mul1;
      } else {
        v[c] = m[c][1];
      }
      t[c][0] = m[c][0];
      t[c][1] = m[c][1];
      t[c][2] = m[c][2];
      m[c][0] = bl;
      m[c][1] = bc;
      m[c][2] = br;
    }
    if (B.keep) {
      for (int c = 0; c < 3; ++c) {
        B.gab[c * plane + size_t(y) * B.ipitch + x] = v[c];
      }
    }
    if (x >= P.xsize || y >= P.ysize) continue;
    const float gamma_r = 1.0f * (v[1] + v[0]);
    const float gamma_g = 1.0f * (v[1] - v[0]);
    const float gamma_b = v[2];
    const float gamma_r2 = gamma_r * gamma_r;
    const float gamma_g2 = gamma_g * gamma_g;
    const float gamma_b2 = gamma_b * gamma_b;
    const float mixed_r = gamma_r2 * gamma_r + P.neg_bias[0];
    const float mixed_g = gamma_g2 * gamma_g + P.neg_bias[1];
    const float mixed_b = gamma_b2 * gamma_b + P.neg_bias[2];
    float lr = inv[0] * mixed_r;
    float lg = inv[3] * mixed_r;
    float lb = inv[6] * mixed_r;
    const float tr = inv[1] * mixed_g;
    const float tg = inv[4] * mixed_g;
    const float tb = inv[7] * mixed_g;
    lr = inv[2] * mixed_b + lr;
    lg = inv[5] * mixed_b + lg;
    lb = inv[8] * mixed_b + lb;
    lr = lr + tr;
    lg = lg + tg;
    lb = lb + tb;
    if (P.grayscale) {
      // kGrayscaleOpt frames: every channel set to the luma of linear RGB,
      // evaluated in double (pik/pik_frame.cc:1155-1166)
      const float gray = (float)(double(lr) * 0.299 + double(lg) * 0.587 +
                                 double(lb) * 0.114);
      lr = lg = lb = gray;
    }
    if (P.out_flags & kOutLinear) {
      const size_t oplane = size_t(P.ysize) * B.opitch;
      B.out[size_t(y) * B.opitch + x] = lr;
      B.out[oplane + size_t(y) * B.opitch + x] = lg;
      B.out[2 * oplane + size_t(y) * B.opitch + x] = lb;
    }
    if (P.out_flags & kOutSrgb8) {
      uint8_t* o = B.srgb8 + size_t(y) * B.spitch + 3 * x;
      o[0] = Srgb8(lr);
      o[1] = Srgb8(lg);
      o[2] = Srgb8(lb);
    }
  }
}
__global__ void __launch_bounds__(256)
    KFinal(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  KFinalBody(P, B, K, PlainBlock());
}

__device__ __forceinline__ void KCopyPlanesBody(
    const float* src, float* dst, size_t n, const VBlock& vb) {
  const size_t i = size_t(vb.x) * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void __launch_bounds__(256)
    KCopyPlanes(const float* src, float* dst, size_t n) {
  KCopyPlanesBody(src, dst, n, PlainBlock());
}

template __global__ void KEpf<true>(Img, FrameParams, FrameBuffers,
                                    const DeviceConsts*, const float*, float*,
                                    int, size_t);
template __global__ void KEpf<false>(Img, FrameParams, FrameBuffers,
                                     const DeviceConsts*, const float*, float*,
                                     int, size_t);

// ---------------------------------------------------------------------------
// Batched forms: one launch over every frame of a batch (BatchSpan), each
// frame's images and buffers as the plain launches of pikgpu_run take them.
__global__ void __launch_bounds__(256)
    KMinMaxPartialDcBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KMinMaxPartialBody(DcImage(P, B), 0, P, B.mm_partial, vb));
}
__global__ void __launch_bounds__(256)
    KMinMaxPartial4Batch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KMinMaxPartial4Body(IdctImage(P, B), s.arg[0], P,
                                        B.mm_partial, vb));
}
__global__ void __launch_bounds__(256) KMinMaxFinalBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KMinMaxFinalBody(B.mm_partial, s.arg[0], B.mm, vb));
}
__global__ void __launch_bounds__(256, 3) KEpfDcBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_ONE(KEpfBody<true>(DcImage(P, B), P, B, K, B.mm, B.dc, P.bx,
                                   size_t(P.bx) * P.by, vb));
}
__global__ void __launch_bounds__(256) KGuideBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_ONE(KGuideBody(IdctImage(P, B), s.arg[0], P, B, B.mm, vb));
}
__global__ void __launch_bounds__(256) KGuideRingBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KGuideRingBody(P, B, vb));
}
__global__ void __launch_bounds__(256, 4) KEpfBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_ONE(KEpfBody<false>(IdctImage(P, B), P, B, K, B.mm, B.ar,
                                    B.ipitch, size_t(B.ipitch) * P.py, vb));
}
__global__ void __launch_bounds__(256) KCopyPlanesBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KCopyPlanesBody(B.idct, B.ar, 3 * size_t(B.ipitch) * P.py,
                                    vb));
}
__global__ void __launch_bounds__(256) KDotsBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KDotsBody(P, B, vb));
}
__global__ void __launch_bounds__(256) KFinalBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KFinalBody(P, B, K, vb));
}

}  // namespace pikgpu
