// DC and LF-domain kernels of the reconstruction (gfx950).
//
// These run at DC resolution (1/64 of the pixels) or on the 2x-upsampled DC
// grid, per 512x512 decode group, and produce the per-group caches the block
// kernel consumes:
//   DequantDC                         pik/compressed_dc.cc:181-230
//   InitializeDecCache (DC + border)  pik/compressed_dc.cc:530-556
//   ComputeLlf                        pik/ac_predictions.cc:380-405
//   PredictLf                         pik/ac_predictions.cc:410-470
//   border of pred2x2                 pik/compressed_image.cc:980-1039
//   UpdateLfForDecoder (pred2x2)      pik/ac_predictions.cc:560-635
// Every kernel is a flat grid over (group, plane, cache position); the group
// caches have fixed pitches (device.h).
#include "device.h"
#include "kernels.h"

namespace pikgpu {

__global__ void __launch_bounds__(256)
    KDcDequant(FrameParams P, FrameBuffers B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.bx * P.by;
  if (i >= n) return;
  // Y first, then X/B = ytox|ytob * Y + q * mul (the portable target's
  // mul_add is a separate multiply and add).
  const float y = float(B.dc16[n + i]) * P.mul_dc[1];
  const float dqx = float(B.dc16[i]) * P.mul_dc[0];
  const float dqb = float(B.dc16[2 * n + i]) * P.mul_dc[2];
  const float x = P.ytox_dc * y + dqx;
  const float b = P.ytob_dc * y + dqb;
  B.dcq[i] = x;
  B.dcq[n + i] = y;
  B.dcq[2 * n + i] = b;
}

// Group geometry helpers.
struct GroupGeom {
  int gbx0, gby0, gw, gh;
};
__device__ inline GroupGeom Geom(const FrameParams& P, int g) {
  GroupGeom G;
  const int gxi = g % P.gx, gyi = g / P.gx;
  G.gbx0 = gxi * kGroupDimBlocks;
  G.gby0 = gyi * kGroupDimBlocks;
  G.gw = min(kGroupDimBlocks, P.bx - G.gbx0);
  G.gh = min(kGroupDimBlocks, P.by - G.gby0);
  return G;
}

// DC of the group cache at cache coordinates (x, y): SourceCoord clamping
// (pik/compressed_dc.h:46-49) is clamp(c - 1, 0, size - 1).
__device__ inline float GroupDc(const FrameParams& P, const float* dcp,
                                const GroupGeom& G, int x, int y) {
  const int sx = min(max(G.gbx0 + x - 1, 0), P.bx - 1);
  const int sy = min(max(G.gby0 + y - 1, 0), P.by - 1);
  return dcp[sy * P.bx + sx];
}

// Strategy of cache position (x, y); border positions count as DCT8 first
// blocks (pik/ac_predictions.cc:391-399).
__device__ inline uint8_t CacheAcs(const FrameParams& P, const FrameBuffers& B,
                                   const GroupGeom& G, int x, int y) {
  if (x == 0 || y == 0 || x == G.gw + 1 || y == G.gh + 1) return 0;
  return B.acs[(G.gby0 + y - 1) * P.bx + G.gbx0 + x - 1];
}

// ---------------------------------------------------------------------------
// ComputeLlf: DC -> lowest-frequency coefficients per cache position.
__global__ void __launch_bounds__(256)
    KComputeLlf(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, c = blockIdx.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  if ((a & 15) != 0) return;  // written by the first block
  const int raw = a >> 4;
  const float* dcp = B.dc + c * P.bx * P.by;
  float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  const int s = AcsCovered(raw);
  if (s == 1) {
    llf[y * kLlfPitch + x] = GroupDc(P, dcp, G, x, y);
    return;
  }
  // ReinterpretingDCT<8s, s, s> of the s x s DC values
  // (pik/ac_strategy.cc:272-293).
  float in[16], out[16];
  for (int iy = 0; iy < s; ++iy) {
    for (int ix = 0; ix < s; ++ix) in[iy * s + ix] = GroupDc(P, dcp, G, x + ix, y + iy);
  }
  SmallDct(K, in, s, out, s, s);
  const int ni = s == 2 ? 0 : 1;  // tscale index of N = s
  const int di = s == 2 ? 3 : 4;  // ascale index of DCTN = 8s
  for (int iy = 0; iy < s; ++iy) {
    for (int ix = 0; ix < s; ++ix) {
      const float v = out[iy * s + ix] * K->tscale[ni][iy][ix];
      llf[(y + iy) * kLlfPitch + x + ix] = v * K->ascale[di][iy][ix];
    }
  }
}

// DC2x2FromLowestFrequencies into tmp2x2 (pik/ac_strategy.cc:928-969).
__global__ void __launch_bounds__(256)
    KDc2x2FromLlf(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                  int to_pred2x2) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, c = blockIdx.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  // Without LF prediction only the border is wanted (the interior of
  // pred2x2 is filled by KPred2x2Interior).
  const bool border = x == 0 || y == 0 || x == W - 1 || y == H - 1;
  if (to_pred2x2 && !border) return;
  if ((a & 15) != 0) return;
  const int raw = a >> 4;
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  float* dst = (to_pred2x2 ? B.pred2x2 : B.tmp2x2) + (size_t(g) * 3 + c) * kG2Plane;
  const int s = AcsCovered(raw);
  float* o = dst + 2 * y * kG2Pitch + 2 * x;
  const float l0 = llf[y * kLlfPitch + x];
  if (s == 1) {
    float v;
    if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
      // ReinterpretingIDCT<8, 1, 2>: one coefficient, 2-point IDCT of it.
      v = (l0 * K->ascale[0][0][0]) * K->tscale[2][0][0];
    } else {
      v = l0;
    }
    o[0] = v;
    o[1] = v;
    o[kG2Pitch] = v;
    o[kG2Pitch + 1] = v;
    return;
  }
  // ReinterpretingIDCT<8s, s, 2s>.
  const int n = 2 * s;
  float blk[64], out[64];
  for (int i = 0; i < n * n; ++i) blk[i] = 0.0f;
  const int ai = s == 2 ? 1 : 2;  // ascale index of N = 2s
  const int ti = s == 2 ? 3 : 4;  // tscale index of DCTN = 8s
  for (int iy = 0; iy < s; ++iy) {
    for (int ix = 0; ix < s; ++ix) {
      const float v = llf[(y + iy) * kLlfPitch + x + ix] * K->ascale[ai][iy][ix];
      blk[iy * n + ix] = v * K->tscale[ti][iy][ix];
    }
  }
  SmallIdct(K, blk, n, out, n, n);
  for (int iy = 0; iy < n; ++iy) {
    for (int ix = 0; ix < n; ++ix) o[iy * kG2Pitch + ix] = out[iy * n + ix];
  }
}

// 3x3 LF prediction blur with mirrored borders (Symmetric3,
// pik/convolve.h:502-640).
__global__ void __launch_bounds__(256)
    KLfBlur(FrameParams P, FrameBuffers B) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, c = blockIdx.z;
  const GroupGeom G = Geom(P, g);
  const int W = 2 * (G.gw + 2), H = 2 * (G.gh + 2);
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const float* in = B.tmp2x2 + (size_t(g) * 3 + c) * kG2Plane;
  float* out = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
  out[y * kG2Pitch + x] =
      Symmetric3(in, kG2Pitch, W, H, x, y, P.lfblur_w[0], P.lfblur_w[1],
                 P.lfblur_w[2]);
}

// LowFrequenciesFromDC2x2, in place on lf2x2 (pik/ac_strategy.cc:1019-1064).
__global__ void __launch_bounds__(256)
    KLfFromDc2x2(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, c = blockIdx.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  if ((a & 15) != 0) return;
  const int raw = a >> 4;
  float* p = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
  const int s = AcsCovered(raw);
  if (s == 1 && !(raw == kDCT || raw == kDCTNoHF || raw == kLines)) {
    const float b00 = p[0], b01 = p[1], b10 = p[kG2Pitch], b11 = p[kG2Pitch + 1];
    p[0] = (b00 + b01 + b10 + b11) * 0.25f;
    p[1] = (b00 + b01 - b10 - b11) * 0.25f;
    p[kG2Pitch] = (b00 - b01 + b10 - b11) * 0.25f;
    p[kG2Pitch + 1] = (b00 - b01 - b10 + b11) * 0.25f;
    return;
  }
  // ReinterpretingDCT<8s, 2s, 2s>.
  const int n = 2 * s;
  float in[64], out[64];
  for (int iy = 0; iy < n; ++iy) {
    for (int ix = 0; ix < n; ++ix) in[iy * n + ix] = p[iy * kG2Pitch + ix];
  }
  SmallDct(K, in, n, out, n, n);
  const int ni = s == 1 ? 0 : (s == 2 ? 1 : 2);  // tscale index of N = 2s
  const int di = s == 1 ? 2 : (s == 2 ? 3 : 4);  // ascale index of 8s
  for (int iy = 0; iy < n; ++iy) {
    for (int ix = 0; ix < n; ++ix) {
      const float v = out[iy * n + ix] * K->tscale[ni][iy][ix];
      p[iy * kG2Pitch + ix] = v * K->ascale[di][iy][ix];
    }
  }
}

// Zero-fills pred2x2 (ZeroFillImage, pik/compressed_image.cc:976).
__global__ void __launch_bounds__(256)
    KZeroPred(FrameParams P, FrameBuffers B) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t n = size_t(P.gx) * P.gy * 3 * kG2Plane;
  if (i < n) B.pred2x2[i] = 0.0f;
}

// Border of pred2x2 from llf and lf2x2 as DCT8 blocks
// (pik/compressed_image.cc:980-1039).
__global__ void __launch_bounds__(256)
    KPred2x2Border(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, c = blockIdx.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  if (!(x == 0 || y == 0 || x == W - 1 || y == H - 1)) return;
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
  float* o = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
  // ReinterpretingIDCT<8, 2, 2> of {llf, lf(0,1), lf(1,0), lf(1,1)}.
  const float a00 = (llf[y * kLlfPitch + x] * K->ascale[0][0][0]) * K->tscale[2][0][0];
  const float a01 = (lf[2 * y * kG2Pitch + 2 * x + 1] * K->ascale[0][0][1]) * K->tscale[2][0][1];
  const float a10 = (lf[(2 * y + 1) * kG2Pitch + 2 * x] * K->ascale[0][1][0]) * K->tscale[2][1][0];
  const float a11 = (lf[(2 * y + 1) * kG2Pitch + 2 * x + 1] * K->ascale[0][1][1]) * K->tscale[2][1][1];
  o[0] = a00 + a01 + a10 + a11;
  o[1] = a00 + a01 - a10 - a11;
  o[kG2Pitch] = a00 - a01 + a10 - a11;
  o[kG2Pitch + 1] = a00 - a01 - a10 + a11;
}

// Interior of pred2x2: per first block, the decoder-side LF coefficients
// (dequantised + LLF + LF prediction) reinterpreted as a 2x upsampled DC
// block (DC2x2FromLowFrequencies, pik/ac_strategy.cc:971-1017, as called by
// UpdateLfForDecoder, pik/ac_predictions.cc:621-633).
__global__ void __launch_bounds__(256)
    KPred2x2Interior(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  const int blk = blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= P.bx * P.by) return;
  const int bxa = blk % P.bx, bya = blk / P.bx;
  const uint8_t a = B.acs[blk];
  if ((a & 15) != 0) return;
  const int raw = a >> 4;
  const int s = AcsCovered(raw);
  const int n = 2 * s;
  const int gxi = bxa / kGroupDimBlocks, gyi = bya / kGroupDimBlocks;
  const int g = gyi * P.gx + gxi;
  const int bxg = bxa - gxi * kGroupDimBlocks, byg = bya - gyi * kGroupDimBlocks;
  float yv[64];
  for (int ci = 0; ci < 3; ++ci) {
    const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);  // Y first for CfL
    const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
    const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
    float F[64];
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        float v;
        if (i < s && j < s) {
          v = llf[(byg + 1 + i) * kLlfPitch + bxg + 1 + j];
        } else {
          const int flat = i * 8 * s + j;
          const int row = bya + flat / (64 * s);
          const int col = bxa * 64 + flat % (64 * s);
          const float y = c == 1 ? 0.0f : yv[i * n + j];
          v = DequantCoef(P, B, c, row, col, y);
          if (P.predict_lf) v += lf[(2 * (byg + 1) + i) * kG2Pitch + 2 * (bxg + 1) + j];
        }
        F[i * n + j] = v;
        if (c == 1 && !(i < s && j < s)) {
          const int flat = i * 8 * s + j;
          yv[i * n + j] = DequantCoef(P, B, 1, bya + flat / (64 * s),
                                      bxa * 64 + flat % (64 * s), 0.0f);
        }
      }
    }
    float* o = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane +
               2 * (byg + 1) * kG2Pitch + 2 * (bxg + 1);
    if (s == 1 && !(raw == kDCT || raw == kDCTNoHF || raw == kLines)) {
      const float b00 = F[0], b01 = F[1], b10 = F[2], b11 = F[3];
      o[0] = b00 + b01 + b10 + b11;
      o[1] = b00 + b01 - b10 - b11;
      o[kG2Pitch] = b00 - b01 + b10 - b11;
      o[kG2Pitch + 1] = b00 - b01 - b10 + b11;
      continue;
    }
    // ReinterpretingIDCT<8s, 2s, 2s>.
    const int ai = s == 1 ? 0 : (s == 2 ? 1 : 2);
    const int ti = s == 1 ? 2 : (s == 2 ? 3 : 4);
    float blkv[64], out[64];
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        const float v = F[i * n + j] * K->ascale[ai][i][j];
        blkv[i * n + j] = v * K->tscale[ti][i][j];
      }
    }
    SmallIdct(K, blkv, n, out, n, n);
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) o[i * kG2Pitch + j] = out[i * n + j];
    }
  }
}

}  // namespace pikgpu
