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

// Also writes each block's dequantisation descriptor (BlockDesc), so the
// block kernels reach the dequant matrices with one dependent load instead
// of the chain strategy -> quant field -> control field -> table index ->
// table offset.
__device__ __forceinline__ void KDcDequantBody(
    FrameParams P, FrameBuffers B, const VBlock& vb) {
  const int i = vb.x * blockDim.x + threadIdx.x;
  const int n = P.bx * P.by;
  if (i >= n) return;
  {
    const int bxa = i % P.bx, bya = i / P.bx;
    const uint8_t a = B.acs[i];
    const int raw = a >> 4, idx = a & 15;
    const int s = AcsCovered(raw);
    const int kind = AcsQuantKind(raw);
    const int obxa = bxa - (idx & (s - 1)), obya = bya - (idx >> (s >> 1));
    const int oblk = obya * P.bx + obxa;
    const int qf = B.qf[i], qfo = B.qf[oblk];
    const int qt = B.dqmap[B.dcf[(bya >> 3) * P.tx + (bxa >> 3)] * 256 + qf - 1];
    const int qto =
        B.dqmap[B.dcf[(obya >> 3) * P.tx + (obxa >> 3)] * 256 + qfo - 1];
    const uint32_t* off = B.offsets + (qt * kNumQuantKinds + kind) * 3;
    const uint32_t* offo = B.offsets + (qto * kNumQuantKinds + kind) * 3;
    B.desc[2 * i] = make_uint4(
        off[0] + idx * 64, off[1] + idx * 64, off[2] + idx * 64,
        __float_as_uint(qf == 0 ? 1E10f : P.inv_global_scale / float(qf)));
    B.desc[2 * i + 1] = make_uint4(
        offo[0] + idx * 64, offo[1] + idx * 64, offo[2] + idx * 64,
        __float_as_uint(P.inv_global_scale / float(qfo)));
  }
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
__global__ void __launch_bounds__(256)
    KDcDequant(FrameParams P, FrameBuffers B) {
  KDcDequantBody(P, B, PlainBlock());
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

// DC of plane c of the group cache at cache coordinates (x, y): SourceCoord
// clamping (pik/compressed_dc.h:46-49) is clamp(c - 1, 0, size - 1). For a
// grayscale frame X and B come from Y (GrayXyb::YToXyb, pik/compressed_image.cc:
// 99-104, in double as there: the piece index truncates (y - sub) * mul * 16,
// the value y * slope + constant is rounded to float), as RestoreXB rebuilds
// them in the group cache before the LF predictions (:953-956).
__device__ inline float GroupDc(const FrameParams& P, const DeviceConsts* K,
                                const float* dc, int c, const GroupGeom& G,
                                int x, int y) {
  const int sx = min(max(G.gbx0 + x - 1, 0), P.bx - 1);
  const int sy = min(max(G.gby0 + y - 1, 0), P.by - 1);
  const size_t plane = size_t(P.bx) * P.by, at = size_t(sy) * P.bx + sx;
  if (!P.grayscale || c == 1) return dc[c * plane + at];
  const float yv = dc[plane + at];
  const GrayXybTable& T = K->gray;
  int i = (int)((yv - T.sub) * T.mul * kGrayPieces);
  i = min(max(0, i), kGrayPieces - 1);
  return c == 0 ? (float)(yv * T.x_slope[i] + T.x_const[i])
                : (float)(yv * T.b_slope[i] + T.b_const[i]);
}

// Strategy of cache position (x, y); border positions count as DCT8 first
// blocks (pik/ac_predictions.cc:391-399).
__device__ inline uint8_t CacheAcs(const FrameParams& P, const FrameBuffers& B,
                                   const GroupGeom& G, int x, int y) {
  if (x == 0 || y == 0 || x == G.gw + 1 || y == G.gh + 1) return 0;
  return B.acs[(G.gby0 + y - 1) * P.bx + G.gbx0 + x - 1];
}

// ---------------------------------------------------------------------------
// ComputeLlf: DC -> lowest-frequency coefficients of the first block at
// cache position (x, y) (LowestFrequenciesFromDC, pik/ac_strategy.cc:
// 850-887).
__device__ inline void LlfAt(const FrameParams& P, const FrameBuffers& B,
                             const DeviceConsts* K, const GroupGeom& G, int g,
                             int c, int x, int y, int raw) {
  float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane + y * kLlfPitch + x;
  if (raw == kDCT16 || raw == kDCT32) {
    // ReinterpretingDCT<8s, s, s> of the s x s DC values.
    float dc[16];
    const int s = raw == kDCT16 ? 2 : 4;
    for (int iy = 0; iy < 4; ++iy) {
      for (int ix = 0; ix < 4; ++ix) {
        dc[iy * 4 + ix] =
            (iy < s && ix < s) ? GroupDc(P, K, B.dc, c, G, x + ix, y + iy) : 0.0f;
      }
    }
    if (raw == kDCT16) {
      const float a00 = dc[0], a01 = dc[1], a10 = dc[4], a11 = dc[5];
      const float d[4] = {(a00 + a01 + a10 + a11) * 0.25f,
                          (a00 + a01 - a10 - a11) * 0.25f,
                          (a00 - a01 + a10 - a11) * 0.25f,
                          (a00 - a01 - a10 + a11) * 0.25f};
      for (int iy = 0; iy < 2; ++iy) {
        for (int ix = 0; ix < 2; ++ix) {
          const float v = d[iy * 2 + ix] * K->tscale[0][iy][ix];
          llf[iy * kLlfPitch + ix] = v * K->ascale[3][iy][ix];
        }
      }
    } else {
      ReinterpDct<32, 4>(K, dc, 4, llf, kLlfPitch);
    }
    return;
  }
  llf[0] = GroupDc(P, K, B.dc, c, G, x, y);
}

// DC2x2FromLowestFrequencies of the first block at cache position (x, y)
// into tmp2x2 or pred2x2 (pik/ac_strategy.cc:928-969).
__device__ inline void Dc2x2At(const FrameParams& P, const FrameBuffers& B,
                               const DeviceConsts* K, int g, int c, int x,
                               int y, int raw, bool to_pred2x2) {
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane + y * kLlfPitch + x;
  float* o = (to_pred2x2 ? B.pred2x2 : B.tmp2x2) + (size_t(g) * 3 + c) * kG2Plane +
             2 * y * kG2Pitch + 2 * x;
  if (raw == kDCT16) {
    ReinterpIdct<16, 2, 4>(K, llf, kLlfPitch, o, kG2Pitch);
  } else if (raw == kDCT32) {
    ReinterpIdct<32, 4, 8>(K, llf, kLlfPitch, o, kG2Pitch);
  } else {
    float v;
    if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
      // ReinterpretingIDCT<8, 1, 2>: one coefficient, 2-point IDCT of it.
      v = (llf[0] * K->ascale[0][0][0]) * K->tscale[2][0][0];
    } else {
      v = llf[0];
    }
    o[0] = v;
    o[1] = v;
    o[kG2Pitch] = v;
    o[kG2Pitch + 1] = v;
  }
}

// LlfAt and, with dc2x2, Dc2x2At of a DCT16 (S = 2) or DCT32 (S = 4) first
// block on N = 2S lanes, each element with the one-lane form's operations
// in the same order. LLF: for S = 4 the 4x4 ReinterpretingDCT runs its
// column pass on lanes 0..3 into the unit's LDS matrix t, then its row pass
// there, lane l ending with LLF column l; for S = 2 every lane recomputes
// the 2x2 Hadamard and keeps its column. The 2x image: the column pass of
// Idct2D<N> over the zero-padded LLF, lane l on column l, into t, then the
// row pass on column l of t. All threads of the workgroup call (barriers);
// `active` is false past the last unit.
template <int S>
__device__ inline void LlfDc2x2Lanes(const FrameParams& P,
                                     const FrameBuffers& B,
                                     const DeviceConsts* K, bool active, int c,
                                     int blk, int l, bool dc2x2, float* t) {
  constexpr int N = 2 * S;
  int g = 0, x = 0, y = 0;
  GroupGeom G = {0, 0, 0, 0};
  if (active) {
    const int bxa = blk % P.bx, bya = blk / P.bx;
    const int gxi = bxa / kGroupDimBlocks, gyi = bya / kGroupDimBlocks;
    g = gyi * P.gx + gxi;
    x = bxa - gxi * kGroupDimBlocks + 1;
    y = bya - gyi * kGroupDimBlocks + 1;
    G = Geom(P, g);
  }
  float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane + y * kLlfPitch + x;
  float col[S];  // LLF column l (lanes l < S)
  if (S == 4) {
    // ReinterpretingDCT<32, 4, 4> of the 4 x 4 DC values
    float v[4], w[4];
    if (active && l < 4) {
#pragma unroll
      for (int iy = 0; iy < 4; ++iy) v[iy] = GroupDc(P, K, B.dc, c, G, x + l, y + iy);
      dct1d::Dct<4>(v, w);
#pragma unroll
      for (int k = 0; k < 4; ++k) t[l * 5 + k] = w[k];
    }
    __syncthreads();
    if (active && l < 4) {
#pragma unroll
      for (int xx = 0; xx < 4; ++xx) v[xx] = t[xx * 5 + l];
      dct1d::Dct<4>(v, w);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float cf = w[j] * (1.0f / 16);
        const float u = cf * K->tscale[1][j][l];
        col[j] = u * K->ascale[4][j][l];
        llf[j * kLlfPitch + l] = col[j];
      }
    }
    __syncthreads();  // t is reused below
  } else {
    if (active) {
      const float a00 = GroupDc(P, K, B.dc, c, G, x, y);
      const float a01 = GroupDc(P, K, B.dc, c, G, x + 1, y);
      const float a10 = GroupDc(P, K, B.dc, c, G, x, y + 1);
      const float a11 = GroupDc(P, K, B.dc, c, G, x + 1, y + 1);
      const float d[4] = {(a00 + a01 + a10 + a11) * 0.25f,
                          (a00 + a01 - a10 - a11) * 0.25f,
                          (a00 - a01 + a10 - a11) * 0.25f,
                          (a00 - a01 - a10 + a11) * 0.25f};
      if (l < 2) {
#pragma unroll
        for (int iy = 0; iy < 2; ++iy) {
          const float v = d[iy * 2 + l] * K->tscale[0][iy][l];
          col[iy] = v * K->ascale[3][iy][l];
          llf[iy * kLlfPitch + l] = col[iy];
        }
      }
    }
  }
  if (!dc2x2) return;  // the same for the whole launch
  // ReinterpretingIDCT<8S, S, 2S>
  constexpr int ai = ScaleIdx<N>(), ti = ScaleIdx<8 * S>();
  float v[N], w[N];
  if (active) {
#pragma unroll
    for (int r = 0; r < N; ++r) {
      v[r] = (r < S && l < S) ? (col[r < S ? r : 0] * K->ascale[ai][r][l]) *
                                    K->tscale[ti][r][l]
                              : 0.0f;
    }
    dct1d::Idct<N>(v, w);
#pragma unroll
    for (int k = 0; k < N; ++k) t[l * (N + 1) + k] = w[k];
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int cc = 0; cc < N; ++cc) v[cc] = t[cc * (N + 1) + l];
  dct1d::Idct<N>(v, w);
  float* o = B.tmp2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
#pragma unroll
  for (int yy = 0; yy < N; ++yy) o[yy * kG2Pitch + l] = w[yy];
}

// ComputeLlf (and with dc2x2 the LLF's 2x image in tmp2x2, the first step
// of PredictLf). One launch: workgroups [0, wg4) take the DCT32 first
// blocks of pred_list (32 units of block x plane on 8 lanes), [wg4, wg4 +
// wg2) the DCT16 ones (64 units of 4 lanes), the rest one lane per cache
// position of (position block, group, plane); there the thread turns its
// block's LLF values into the 2x image reading back what it wrote.
__device__ __forceinline__ void KComputeLlfBody(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, int dc2x2, int wg4,
    int wg2, int pos_blocks, int ngroups, const VBlock& vb) {
  __shared__ float s_t[32 * 8 * 9];
  const int tid = threadIdx.x;
  const int wg = vb.x;
  if (wg < wg4 + wg2) {
    const bool big = wg < wg4;
    const int unit = big ? wg * 32 + (tid >> 3) : (wg - wg4) * 64 + (tid >> 2);
    const int kb = unit / 3, c = unit - 3 * kb;
    const int first = big ? P.pred_end[2] : P.pred_end[1];
    const int n = big ? P.pred_end[3] - P.pred_end[2]
                      : P.pred_end[2] - P.pred_end[1];
    const bool active = kb < n;
    const int blk = active ? int(B.pred_list[first + kb]) : 0;
    if (big) {
      LlfDc2x2Lanes<4>(P, B, K, active, c, blk, tid & 7, dc2x2 != 0,
                       s_t + (tid >> 3) * 8 * 9);
    } else {
      LlfDc2x2Lanes<2>(P, B, K, active, c, blk, tid & 3, dc2x2 != 0,
                       s_t + (tid >> 2) * 4 * 5);
    }
    return;
  }
  const int r = wg - wg4 - wg2;
  const int pb = r % pos_blocks, g = (r / pos_blocks) % ngroups;
  const int c = r / (pos_blocks * ngroups);
  const int pos = pb * 256 + tid;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  if ((a & 15) != 0) return;  // written by the first block
  const int raw = a >> 4;
  if (raw == kDCT16 || raw == kDCT32) return;  // the lane segments
  LlfAt(P, B, K, G, g, c, x, y, raw);
  if (dc2x2) Dc2x2At(P, B, K, g, c, x, y, raw, false);
}
__global__ void __launch_bounds__(256)
    KComputeLlf(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                int dc2x2, int wg4, int wg2, int pos_blocks, int ngroups) {
  KComputeLlfBody(P, B, K, dc2x2, wg4, wg2, pos_blocks, ngroups, PlainBlock());
}

// The border of pred2x2 from the LLF cache without LF prediction
// (pik/compressed_image.cc:980-1039; the interior is KPred2x2Interior's).
__device__ __forceinline__ void KDc2x2BorderBody(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, const VBlock& vb) {
  const int pos = vb.x * blockDim.x + threadIdx.x;
  const int g = vb.y, c = vb.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  if (!(x == 0 || y == 0 || x == W - 1 || y == H - 1)) return;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  if ((a & 15) != 0) return;
  Dc2x2At(P, B, K, g, c, x, y, a >> 4, true);
}
__global__ void __launch_bounds__(256)
    KDc2x2Border(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  KDc2x2BorderBody(P, B, K, PlainBlock());
}

// 3x3 LF prediction blur with mirrored borders (Symmetric3,
// pik/convolve.h:502-640).
__device__ __forceinline__ void KLfBlurBody(
    FrameParams P, FrameBuffers B, const VBlock& vb) {
  const int pos = vb.x * blockDim.x + threadIdx.x;
  const int g = vb.y, c = vb.z;
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
__global__ void __launch_bounds__(256)
    KLfBlur(FrameParams P, FrameBuffers B) {
  KLfBlurBody(P, B, PlainBlock());
}

// ReinterpDct<8S, 2S> in place on the 2S x 2S block at p (pitch kG2Pitch)
// on N = 2S lanes: lane l runs the first pass of Dct2D over column l into
// row l of the unit's LDS matrix t (pitch N + 1), then, after the barrier,
// the second pass over column l of t, scaling and writing column l. Every
// element sees the one-lane form's operations in the same order. All
// threads of the workgroup call; `active` is false past the last unit.
template <int S>
__device__ inline void ReinterpDctLanes(const DeviceConsts* K, bool active,
                                        float* p, int l, float* t) {
  constexpr int N = 2 * S;
  constexpr int ti = ScaleIdx<N>(), ai = ScaleIdx<8 * S>();
  float v[N], w[N];
  if (active) {
#pragma unroll
    for (int y = 0; y < N; ++y) v[y] = p[y * kG2Pitch + l];
    dct1d::Dct<N>(v, w);
#pragma unroll
    for (int k = 0; k < N; ++k) t[l * (N + 1) + k] = w[k];
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int x = 0; x < N; ++x) v[x] = t[x * (N + 1) + l];
  dct1d::Dct<N>(v, w);
  const float scale = 1.0f / (N * N);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const float cf = w[j] * scale;
    const float u = cf * K->tscale[ti][j][l];
    p[j * kG2Pitch + l] = u * K->ascale[ai][j][l];
  }
}

// LowFrequenciesFromDC2x2, in place on lf2x2 (pik/ac_strategy.cc:1019-1064).
// One launch: workgroups [0, wg4) take the DCT32 first blocks of pred_list
// (32 units of block x plane on 8 lanes each), [wg4, wg4 + wg2) the DCT16
// ones (64 units of 4 lanes), the rest one lane per cache position of
// (position block, group, plane) for the other strategies.
__device__ __forceinline__ void KLfFromDc2x2Body(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, int wg4, int wg2,
    int pos_blocks, int ngroups, const VBlock& vb) {
  __shared__ float s_t[32 * 8 * 9];
  const int tid = threadIdx.x;
  const int wg = vb.x;
  if (wg < wg4 + wg2) {
    const bool big = wg < wg4;
    const int lanes = big ? 8 : 4;
    const int unit = big ? wg * 32 + (tid >> 3) : (wg - wg4) * 64 + (tid >> 2);
    const int kb = unit / 3, c = unit - 3 * kb;
    const int first = big ? P.pred_end[2] : P.pred_end[1];
    const int n = big ? P.pred_end[3] - P.pred_end[2]
                      : P.pred_end[2] - P.pred_end[1];
    const bool active = kb < n;
    float* p = nullptr;
    if (active) {
      const int blk = int(B.pred_list[first + kb]);
      const int bxa = blk % P.bx, bya = blk / P.bx;
      const int gxi = bxa / kGroupDimBlocks, gyi = bya / kGroupDimBlocks;
      const int g = gyi * P.gx + gxi;
      const int x = bxa - gxi * kGroupDimBlocks + 1;
      const int y = bya - gyi * kGroupDimBlocks + 1;
      p = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
    }
    if (big) {
      ReinterpDctLanes<4>(K, active, p, tid & 7, s_t + (tid >> 3) * 8 * 9);
    } else {
      ReinterpDctLanes<2>(K, active, p, tid & 3, s_t + (tid >> 2) * 4 * 5);
    }
    (void)lanes;
    return;
  }
  const int r = wg - wg4 - wg2;
  const int pb = r % pos_blocks, g = (r / pos_blocks) % ngroups;
  const int c = r / (pos_blocks * ngroups);
  const int pos = pb * 256 + tid;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  const uint8_t a = CacheAcs(P, B, G, x, y);
  if ((a & 15) != 0) return;
  const int raw = a >> 4;
  if (raw == kDCT16 || raw == kDCT32) return;  // the lane segments
  float* p = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
  const float b00 = p[0], b01 = p[1], b10 = p[kG2Pitch], b11 = p[kG2Pitch + 1];
  const float d[4] = {(b00 + b01 + b10 + b11) * 0.25f,
                      (b00 + b01 - b10 - b11) * 0.25f,
                      (b00 - b01 + b10 - b11) * 0.25f,
                      (b00 - b01 - b10 + b11) * 0.25f};
  if (raw == kDCT || raw == kDCTNoHF || raw == kLines) {
    // ReinterpretingDCT<8, 2, 2>
    for (int iy = 0; iy < 2; ++iy) {
      for (int ix = 0; ix < 2; ++ix) {
        const float v = d[iy * 2 + ix] * K->tscale[0][iy][ix];
        p[iy * kG2Pitch + ix] = v * K->ascale[2][iy][ix];
      }
    }
  } else {
    p[0] = d[0];
    p[1] = d[1];
    p[kG2Pitch] = d[2];
    p[kG2Pitch + 1] = d[3];
  }
}
__global__ void __launch_bounds__(256)
    KLfFromDc2x2(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                 int wg4, int wg2, int pos_blocks, int ngroups) {
  KLfFromDc2x2Body(P, B, K, wg4, wg2, pos_blocks, ngroups, PlainBlock());
}

// Border of pred2x2 from llf and lf2x2 as DCT8 blocks
// (pik/compressed_image.cc:980-1039).
template <bool kUseLf>
__device__ __forceinline__ void KPred2x2BorderBody(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, const VBlock& vb) {
  const int pos = vb.x * blockDim.x + threadIdx.x;
  const int g = vb.y, c = vb.z;
  const GroupGeom G = Geom(P, g);
  const int W = G.gw + 2, H = G.gh + 2;
  if (pos >= W * H) return;
  const int x = pos % W, y = pos / W;
  if (!(x == 0 || y == 0 || x == W - 1 || y == H - 1)) return;
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
  float* o = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane + 2 * y * kG2Pitch + 2 * x;
  // ReinterpretingIDCT<8, 2, 2> of {llf, lf(0,1), lf(1,0), lf(1,1)}; the
  // encoder without LF prediction forms the same with zeros in place of
  // the LF values (pik/ac_predictions.cc:187-198).
  const float l01 = kUseLf ? lf[2 * y * kG2Pitch + 2 * x + 1] : 0.0f;
  const float l10 = kUseLf ? lf[(2 * y + 1) * kG2Pitch + 2 * x] : 0.0f;
  const float l11 = kUseLf ? lf[(2 * y + 1) * kG2Pitch + 2 * x + 1] : 0.0f;
  const float a00 = (llf[y * kLlfPitch + x] * K->ascale[0][0][0]) * K->tscale[2][0][0];
  const float a01 = (l01 * K->ascale[0][0][1]) * K->tscale[2][0][1];
  const float a10 = (l10 * K->ascale[0][1][0]) * K->tscale[2][1][0];
  const float a11 = (l11 * K->ascale[0][1][1]) * K->tscale[2][1][1];
  o[0] = a00 + a01 + a10 + a11;
  o[1] = a00 + a01 - a10 - a11;
  o[kG2Pitch] = a00 - a01 + a10 - a11;
  o[kG2Pitch + 1] = a00 - a01 - a10 + a11;
}
template <bool kUseLf>
__global__ void __launch_bounds__(256)
    KPred2x2Border(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  KPred2x2BorderBody<kUseLf>(P, B, K, PlainBlock());
}

// Interior of pred2x2: per first block and plane, the decoder-side LF
// coefficients (dequantised + LLF + LF prediction) reinterpreted as a 2x
// upsampled DC block (DC2x2FromLowFrequencies, pik/ac_strategy.cc:971-1017,
// as called by UpdateLfForDecoder, pik/ac_predictions.cc:621-633). X and B
// recompute the dequantised Y values their colour correlation needs.
//
// kEnc: the encoder's side (ComputeDecoderBlockAnd2x2DC, pik/ac_predictions.
// cc:136-199): the corner's values were formed by KEncLf and wait in
// `corner` (coefficient layout).
template <int S, bool kEnc = false>
__device__ inline void Pred2x2Block(const FrameParams& P, const FrameBuffers& B,
                                    const DeviceConsts* K, int c, int bxa,
                                    int bya, int bxg, int byg, int g,
                                    const float* corner) {
  constexpr int N = 2 * S;  // LF corner size
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
  // Dequantisation parameters of the storage blocks the corner touches:
  // coefficient F[i][j] (flat i*8S + j) lies in storage block
  // (bxa + flat / 64, bya) for these sizes.
  constexpr int kStore = S == 4 ? 4 : 1;
  BlockDq dq[kStore], dqy[kStore];
#pragma unroll
  for (int k = 0; k < kStore; ++k) {
    dq[k] = DescBlockDq(P, B, c, bya, bxa + k);
    if (c != 1) dqy[k] = DescBlockDq(P, B, 1, bya, bxa + k);
  }
  const size_t row0 = size_t(bya) * P.bx * 64 + size_t(bxa) * 64;
  const size_t cplane = size_t(c) * P.by * P.bx * 64;
  float F[N * N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      float v;
      if (kEnc) {
        v = corner[cplane + row0 + i * 8 * S + j];
      } else if (i < S && j < S) {
        v = llf[(byg + 1 + i) * kLlfPitch + bxg + 1 + j];
      } else {
        const int flat = i * 8 * S + j;
        const int sb = S == 4 ? flat >> 6 : 0, k = flat & 63;
        v = DequantPasses(P, B, dq[sb], dqy[sb], c, row0 + flat, k);
        if (P.predict_lf) {
          v += lf[(2 * (byg + 1) + i) * kG2Pitch + 2 * (bxg + 1) + j];
        }
      }
      F[i * N + j] = v;
    }
  }
  float* o = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane +
             2 * (byg + 1) * kG2Pitch + 2 * (bxg + 1);
  // ReinterpretingIDCT<8S, 2S, 2S>
  if (S == 1) {
    const float a00 = (F[0] * K->ascale[0][0][0]) * K->tscale[2][0][0];
    const float a01 = (F[1] * K->ascale[0][0][1]) * K->tscale[2][0][1];
    const float a10 = (F[2] * K->ascale[0][1][0]) * K->tscale[2][1][0];
    const float a11 = (F[3] * K->ascale[0][1][1]) * K->tscale[2][1][1];
    Idct2x2(a00, a01, a10, a11, o, kG2Pitch);
  } else {
    ReinterpIdct<8 * S, N, N>(K, F, N, o, kG2Pitch);
  }
}

// Location of a first block in its group.
struct FirstBlock {
  int bxa, bya, bxg, byg, g;
};
__device__ inline FirstBlock LocateBlock(const FrameParams& P, int blk) {
  FirstBlock f;
  f.bxa = blk % P.bx;
  f.bya = blk / P.bx;
  const int gxi = f.bxa / kGroupDimBlocks, gyi = f.bya / kGroupDimBlocks;
  f.g = gyi * P.gx + gxi;
  f.bxg = f.bxa - gxi * kGroupDimBlocks;
  f.byg = f.bya - gyi * kGroupDimBlocks;
  return f;
}

// Pred2x2Block<S> for S = 2, 4 on N = 2S lanes of a workgroup: lane l
// builds column l of the corner F, scales it and runs the column pass of
// Idct2D into row l of the unit's LDS matrix t (pitch N + 1); after the
// barrier lane l runs the row pass over column l of t and writes column l
// of the output. Each element sees the operations of the one-lane form in
// the same order. Every thread of the workgroup calls (the barrier);
// `active` is false for lanes past the last unit.
template <int S, bool kEnc = false>
__device__ inline void Pred2x2Lanes(const FrameParams& P, const FrameBuffers& B,
                                    const DeviceConsts* K, bool active, int c,
                                    int blk, int l, float* t,
                                    const float* corner) {
  constexpr int N = 2 * S;
  constexpr int ai = ScaleIdx<N>(), ti = ScaleIdx<8 * S>();
  FirstBlock f = {0, 0, 0, 0, 0};
  float v[N], w[N];
  if (active) {
    f = LocateBlock(P, blk);
    const float* llf = B.llf + (size_t(f.g) * 3 + c) * kLlfPlane;
    const float* lf = B.lf2x2 + (size_t(f.g) * 3 + c) * kG2Plane;
    const size_t row0 = size_t(f.bya) * P.bx * 64 + size_t(f.bxa) * 64;
    const size_t cplane = size_t(c) * P.by * P.bx * 64;
#pragma unroll
    for (int y = 0; y < N; ++y) {
      float x;
      if (kEnc) {
        x = corner[cplane + row0 + y * 8 * S + l];
      } else if (y < S && l < S) {
        x = llf[(f.byg + 1 + y) * kLlfPitch + f.bxg + 1 + l];
      } else {
        // F[y][l] is flat y * 8S + l, in storage block bxa + flat / 64
        const int flat = y * 8 * S + l;
        const int sb = flat >> 6, k = flat & 63;
        const BlockDq d = DescBlockDq(P, B, c, f.bya, f.bxa + sb);
        BlockDq dy = d;
        if (c != 1) dy = DescBlockDq(P, B, 1, f.bya, f.bxa + sb);
        x = DequantPasses(P, B, d, dy, c, row0 + flat, k);
        if (P.predict_lf) {
          x += lf[(2 * (f.byg + 1) + y) * kG2Pitch + 2 * (f.bxg + 1) + l];
        }
      }
      // ReinterpretingIDCT<8S, 2S, 2S>'s scaling
      v[y] = (x * K->ascale[ai][y][l]) * K->tscale[ti][y][l];
    }
    dct1d::Idct<N>(v, w);
#pragma unroll
    for (int x = 0; x < N; ++x) t[l * (N + 1) + x] = w[x];
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int cc = 0; cc < N; ++cc) v[cc] = t[cc * (N + 1) + l];
  dct1d::Idct<N>(v, w);
  float* o = B.pred2x2 + (size_t(f.g) * 3 + c) * kG2Plane +
             2 * (f.byg + 1) * kG2Pitch + 2 * (f.bxg + 1);
#pragma unroll
  for (int y = 0; y < N; ++y) o[y * kG2Pitch + l] = w[y];
}

// Interior of pred2x2 over the first blocks of pred_list, binned on the
// host by class (DCT8-like, 2x2 Hadamard, DCT16, DCT32; P.pred_end holds
// the class ends). One launch in three segments, the longest work first:
// workgroups [0, wg4) take the DCT32 blocks, 32 units (block x plane) of 8
// lanes; [wg4, wg4 + wg2) the DCT16 blocks, 64 units of 4 lanes; the rest
// one lane per unit for the DCT8-like and Hadamard classes. The three
// planes of a block are adjacent units (X and B read the block's Y
// coefficients and descriptors with the Y unit).
template <bool kEnc>
__device__ __forceinline__ void KPred2x2InteriorBody(
    FrameParams P, FrameBuffers B, const DeviceConsts* K, int wg4, int wg2,
    const float* corner, const VBlock& vb) {
  __shared__ float s_t[32 * 8 * 9];
  const int tid = threadIdx.x;
  const int wg = vb.x;
  if (wg < wg4) {
    const int unit = wg * 32 + (tid >> 3);
    const int kb = unit / 3, c = unit - 3 * kb;
    const bool active = kb < P.pred_end[3] - P.pred_end[2];
    const int blk = active ? int(B.pred_list[P.pred_end[2] + kb]) : 0;
    Pred2x2Lanes<4, kEnc>(P, B, K, active, c, blk, tid & 7,
                          s_t + (tid >> 3) * 8 * 9, corner);
    return;
  }
  if (wg < wg4 + wg2) {
    const int unit = (wg - wg4) * 64 + (tid >> 2);
    const int kb = unit / 3, c = unit - 3 * kb;
    const bool active = kb < P.pred_end[2] - P.pred_end[1];
    const int blk = active ? int(B.pred_list[P.pred_end[1] + kb]) : 0;
    Pred2x2Lanes<2, kEnc>(P, B, K, active, c, blk, tid & 3,
                          s_t + (tid >> 2) * 4 * 5, corner);
    return;
  }
  const int lane = (wg - wg4 - wg2) * 256 + tid;
  const int k = lane / 3, c = lane - 3 * k;
  if (k >= P.pred_end[1]) return;
  const FirstBlock f = LocateBlock(P, int(B.pred_list[k]));
  if (k < P.pred_end[0]) {
    Pred2x2Block<1, kEnc>(P, B, K, c, f.bxa, f.bya, f.bxg, f.byg, f.g,
                          corner);
    return;
  }
  // DCT2X2, DCT4X4(_NOHF), IDENTITY: 2x2 Hadamard of the LF corner.
  const int bxa = f.bxa, bya = f.bya, bxg = f.bxg, byg = f.byg, g = f.g;
  const BlockDq dq = DescBlockDq(P, B, c, bya, bxa);
  BlockDq dqy;
  if (c != 1) dqy = DescBlockDq(P, B, 1, bya, bxa);
  const size_t row0 = size_t(bya) * P.bx * 64 + size_t(bxa) * 64;
  const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
  const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
  float F[4];
  const size_t cplane = size_t(c) * P.by * P.bx * 64;
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j) {
      float v;
      if (kEnc) {
        v = corner[cplane + row0 + i * 8 + j];
      } else if (i == 0 && j == 0) {
        v = llf[(byg + 1) * kLlfPitch + bxg + 1];
      } else {
        const int flat = i * 8 + j;
        v = DequantPasses(P, B, dq, dqy, c, row0 + flat, flat);
        if (P.predict_lf) {
          v += lf[(2 * (byg + 1) + i) * kG2Pitch + 2 * (bxg + 1) + j];
        }
      }
      F[i * 2 + j] = v;
    }
  }
  float* o = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane +
             2 * (byg + 1) * kG2Pitch + 2 * (bxg + 1);
  Idct2x2(F[0], F[1], F[2], F[3], o, kG2Pitch);
}
template <bool kEnc>
__global__ void __launch_bounds__(256)
    KPred2x2Interior(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                     int wg4, int wg2, const float* corner) {
  KPred2x2InteriorBody<kEnc>(P, B, K, wg4, wg2, corner, PlainBlock());
}

template __global__ void KPred2x2Border<true>(FrameParams, FrameBuffers,
                                              const DeviceConsts*);
template __global__ void KPred2x2Border<false>(FrameParams, FrameBuffers,
                                               const DeviceConsts*);
template __global__ void KPred2x2Interior<false>(FrameParams, FrameBuffers,
                                                 const DeviceConsts*, int, int,
                                                 const float*);
template __global__ void KPred2x2Interior<true>(FrameParams, FrameBuffers,
                                                const DeviceConsts*, int, int,
                                                const float*);

// ---------------------------------------------------------------------------
// Batched forms: one launch over every frame of a batch, each frame's part
// of the launch running the plain kernel's grid (BatchSpan, with the
// plain launch's integer arguments in arg[]).
__global__ void __launch_bounds__(256) KDcDequantBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KDcDequantBody(P, B, vb));
}
__global__ void __launch_bounds__(256) KComputeLlfBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KComputeLlfBody(P, B, K, s.arg[0], s.arg[1], s.arg[2],
                                    s.arg[3], s.arg[4], vb));
}
__global__ void __launch_bounds__(256) KDc2x2BorderBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KDc2x2BorderBody(P, B, K, vb));
}
__global__ void __launch_bounds__(256) KLfBlurBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KLfBlurBody(P, B, vb));
}
__global__ void __launch_bounds__(256) KLfFromDc2x2Batch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KLfFromDc2x2Body(P, B, K, s.arg[0], s.arg[1], s.arg[2],
                                     s.arg[3], vb));
}
__global__ void __launch_bounds__(256)
    KPred2x2BorderBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_LOOP(KPred2x2BorderBody<true>(P, B, K, vb));
}
__global__ void __launch_bounds__(256)
    KPred2x2InteriorBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_ONE(KPred2x2InteriorBody<false>(P, B, K, s.arg[0], s.arg[1],
                                                nullptr, vb));
}

}  // namespace pikgpu
