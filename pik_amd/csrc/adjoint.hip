// Forward model and adjoint of the differentiable part of the
// reconstruction, for gradient-based quant-field search (gfx950, wave64).
//
// The model is the decode with HF prediction, adaptive reconstruction and
// dots off, as a function of a real-valued quant field qf (one value per
// block; the dequantisation table stays the one the frame's integer field
// selects):
//   d_c   = AdjustQuantBias_c(q) * M_c[k] * (inv_global_scale / qf)
//           (Dequant::DoAC, pik/compressed_image.cc:586-693)
//   C'    = LLF / LF predictions + CfL(d)   (pik/ac_predictions.cc:560-620)
//   P     = TransformToPixels(C')           (pik/ac_strategy.cc:674-840)
//   G     = Gaborish(P), cropped            (pik/compressed_image.cc:1164)
//   L     = OpsinToLinear(G)                (pik/opsin_inverse.cc:39-123)
// With qf equal to the integer field, L is bit for bit the decoder's output
// for the same flags.
//
// Two kernels per evaluation, one 32x32-pixel tile per workgroup:
//   KAdjForward   dequant (real qf) + LLF/LF + CfL + IDCT -> P in HBM
//   KAdjBackward  P with a 2-pixel ring -> G, L, dL/dL (from a target or
//                 given), opsin adjoint -> dG with a 1-pixel ring (mirrored
//                 at the image edge), Gaborish transpose -> dP, transform
//                 transposes -> dC', CfL transpose -> dd, and
//                 dL/dqf = -sum_k dd . d / qf per block.
// P is written once and read once (plus ring); nothing else is stored
// between the two kernels.
#include "kernels.h"

namespace pikgpu {

namespace {

// Coefficient ownership of a 256-thread tile pass, as in KBlocks: storage
// row crow of the tile, columns ccol..ccol+3; 16 consecutive lanes own one
// 8x8 storage block.
struct CoefSlot {
  int crow, ccol, cbx, cby;
  bool valid;
  int raw, cs, cN, lgn, cflat0, obxg, obyg;
  uint8_t acs;
};

__device__ inline CoefSlot MakeSlot(const FrameParams& P, const uint8_t* s_acs,
                                    int tbx, int tby, int gbx0, int gby0) {
  CoefSlot s;
  const int tid = threadIdx.x;
  s.crow = tid >> 6;
  s.ccol = (tid & 63) << 2;
  s.cbx = tbx + (s.ccol >> 6);
  s.cby = tby + s.crow;
  s.valid = s.cbx < P.bx && s.cby < P.by;
  s.acs = s_acs[s.crow * 4 + (s.ccol >> 6)];
  s.raw = s.acs >> 4;
  const int cidx = s.acs & 15;
  s.cs = s.acs == 0xFF ? 1 : AcsCovered(s.raw);
  const int lgs = s.cs >> 1;  // log2(cs): cs is 1, 2 or 4
  s.lgn = 3 + lgs;
  const int coby = s.crow - (cidx >> lgs), cobx = (s.ccol >> 6) - (cidx & (s.cs - 1));
  s.cflat0 = (s.crow - coby) * 64 * s.cs + (s.ccol - cobx * 64);
  s.cN = 8 * s.cs;
  s.obxg = tbx + cobx - gbx0;
  s.obyg = tby + coby - gby0;
  return s;
}

__device__ inline void LoadTileAcs(const FrameParams& P, const FrameBuffers& B,
                                   int tbx, int tby, uint8_t* s_acs) {
  const int tid = threadIdx.x;
  if (tid < 16) {
    const int bxa = tbx + (tid & 3), bya = tby + (tid >> 2);
    s_acs[tid] = (bxa < P.bx && bya < P.by) ? B.acs[bya * P.bx + bxa] : 0xFF;
  }
  __syncthreads();
}

// Dequantised AC value of channel c without colour correlation, in
// DequantWith's operation order with the real-valued field's scale.
__device__ inline float DequantPlain(const float* m, float scaled, int c,
                                     int16_t q, int k) {
  const float mul = m[k] * scaled;
  return AdjustQuantBias(c, float(q)) * mul;
}

}  // namespace

__global__ void __launch_bounds__(256)
    KAdjForward(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                AdjArgs A) {
  __shared__ float s_coef[3][kCoefPlane];
  __shared__ float s_pix[3][kTilePlane];
  __shared__ float s_tmp[3][kTilePlane];
  __shared__ uint8_t s_acs[16];
  __shared__ TileTasks s_inv;

  const int tiles_x = (P.bx + 3) / 4;
  const int tile = XcdTile(blockIdx.x, tiles_x * ((P.by + 3) / 4));
  if (tile < 0) return;
  const int tbx = (tile % tiles_x) * 4, tby = (tile / tiles_x) * 4;
  LoadTileAcs(P, B, tbx, tby, s_acs);
  BuildTasks(s_acs, false, &s_inv);

  const int gxi = tbx / kGroupDimBlocks, gyi = tby / kGroupDimBlocks;
  const int g = gyi * P.gx + gxi;
  const CoefSlot S = MakeSlot(P, s_acs, tbx, tby, gxi * kGroupDimBlocks,
                              gyi * kGroupDimBlocks);
  if (S.valid) {
    const size_t plane = size_t(P.by) * P.bx * 64;
    const size_t at =
        size_t(S.cby) * P.bx * 64 + size_t(S.cbx) * 64 + (S.ccol & 63);
    const float scaled =
        P.inv_global_scale / A.qf[size_t(S.cby) * A.qf_pitch + S.cbx];
    float yv[4];
    for (int ci = 0; ci < 3; ++ci) {
      const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
      const BlockDq dq = MakeBlockDq(P, B, c, S.cby, S.cbx);
      const int16_t* q = B.ac + c * plane + at;
      const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
      const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
      for (int j = 0; j < 4; ++j) {
        const float v0 = DequantPlain(dq.m, scaled, c, q[j], (S.ccol & 63) + j);
        float v = c == 1 ? v0 : dq.cc * yv[j] + v0;
        if (c == 1) yv[j] = v;
        const int flat = S.cflat0 + j;
        const int fi = flat >> S.lgn, fj = flat & (S.cN - 1);
        if (fi < S.cs && fj < S.cs) {
          v = llf[(S.obyg + 1 + fi) * kLlfPitch + S.obxg + 1 + fj];
        } else if (P.predict_lf && fi < 2 * S.cs && fj < 2 * S.cs) {
          v += lf[(2 * (S.obyg + 1) + fi) * kG2Pitch + 2 * (S.obxg + 1) + fj];
        }
        s_coef[c][CoefAt(S.crow, S.ccol + j)] = v;
      }
    }
  }
  __syncthreads();
  TileInverse3(s_inv, &s_coef[0][0], &s_pix[0][0], &s_tmp[0][0]);
  const int tid = threadIdx.x;
  const int pr = tid >> 3, pc0 = (tid & 7) << 2;
  const int py = tby * 8 + pr;
  if (py < P.py) {
    for (int c = 0; c < 3; ++c) {
      float* dst = B.idct + (size_t(c) * P.py + py) * B.ipitch + tbx * 8 + pc0;
      for (int j = 0; j < 4; ++j) {
        if (tbx * 8 + pc0 + j < P.px) dst[j] = s_pix[c][PixAt(pr, pc0 + j)];
      }
    }
  }
}

// Ring geometry of the backward tile: P at [-2, 34) and dG at [-1, 33)
// around the 32x32 tile.
constexpr int kPW = 36;
constexpr int kGW = 34;

__global__ void __launch_bounds__(256)
    KAdjBackward(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                 AdjArgs A) {
  // s_a (the P ring) is dead once dG is formed; the forward transform's
  // scratch reuses it. s_b (the dG ring) is dead once dP is formed; the
  // coefficients reuse it.
  static_assert(kPW * kPW >= kTilePlane && kGW * kGW >= kCoefPlane,
                "the rings must hold the transform buffers");
  __shared__ float s_a[3 * kPW * kPW];  // P ring, then transform scratch
  __shared__ float s_b[3 * kGW * kGW];  // dG ring, then coefficients
  __shared__ float s_pix[3][kTilePlane];  // dP of the tile
  __shared__ uint8_t s_acs[16];
  __shared__ TileTasks s_fwd;
  __shared__ float s_loss[4];

  const int tid = threadIdx.x;
  const int tiles_x = (P.bx + 3) / 4;
  const int num_tiles = tiles_x * ((P.by + 3) / 4);
  const int tile = XcdTile(blockIdx.x, num_tiles);
  if (tile < 0) return;
  const int tbx = (tile % tiles_x) * 4, tby = (tile / tiles_x) * 4;
  const int ox = tbx * 8, oy = tby * 8;
  LoadTileAcs(P, B, tbx, tby, s_acs);
  BuildTasks(s_acs, false, &s_fwd);

  // 1. P ring (mirrored at the 8-padded image edge, as the Gaborish
  //    convolution reads it).
  const size_t iplane = size_t(P.py) * B.ipitch;
  for (int k = tid; k < kPW * kPW; k += 256) {
    const int Y = Mirror(oy - 2 + k / kPW, P.py);
    const int X = Mirror(ox - 2 + k % kPW, P.px);
    const float* src = B.idct + size_t(Y) * B.ipitch + X;
    for (int c = 0; c < 3; ++c) s_a[c * kPW * kPW + k] = src[c * iplane];
  }
  __syncthreads();

  // 2. G, L and dL/dG at the ring positions inside the padded image.
  const float* inv = K->opsin_inv;
  float loss = 0.0f;
  for (int k = tid; k < kGW * kGW; k += 256) {
    const int i = k / kGW, j = k % kGW;
    const int Y = oy - 1 + i, X = ox - 1 + j;
    if (Y < 0 || Y >= P.py || X < 0 || X >= P.px) continue;
    float v[3];
    for (int c = 0; c < 3; ++c) {
      const float* pc = s_a + c * kPW * kPW;
      v[c] = P.gaborish ? Symmetric3(pc, kPW, kPW, kPW, j + 1, i + 1,
                                     P.gab_w[0], P.gab_w[1], P.gab_w[2])
                        : pc[(i + 1) * kPW + j + 1];
    }
    float dv0 = 0.0f, dv1 = 0.0f, dv2 = 0.0f;
    if (Y < P.ysize && X < P.xsize) {
      // OpsinToLinear in KFinal's operation order.
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
      const bool inner = i >= 1 && i <= 32 && j >= 1 && j <= 32;
      const size_t at = size_t(Y) * A.in_pitch + X;
      float e0, e1, e2;
      if (A.target[0]) {
        e0 = lr - A.target[0][at];
        e1 = lg - A.target[1][at];
        e2 = lb - A.target[2][at];
        if (inner) loss += e0 * e0 + e1 * e1 + e2 * e2;
        e0 *= 2.0f;
        e1 *= 2.0f;
        e2 *= 2.0f;
      } else {
        e0 = A.upstream[0][at];
        e1 = A.upstream[1][at];
        e2 = A.upstream[2][at];
      }
      if (inner && A.linear[0]) {
        const size_t o = size_t(Y) * A.lin_pitch + X;
        A.linear[0][o] = lr;
        A.linear[1][o] = lg;
        A.linear[2][o] = lb;
      }
      // Opsin transpose: dmixed_j = sum_i inv[3i + j] dL_i; d gamma =
      // 3 gamma^2 dmixed; (r, g) = (Y + X, Y - X).
      const float dmr = inv[0] * e0 + inv[3] * e1 + inv[6] * e2;
      const float dmg = inv[1] * e0 + inv[4] * e1 + inv[7] * e2;
      const float dmb = inv[2] * e0 + inv[5] * e1 + inv[8] * e2;
      const float dgr = 3.0f * gamma_r2 * dmr;
      const float dgg = 3.0f * gamma_g2 * dmg;
      dv0 = dgr - dgg;
      dv1 = dgr + dgg;
      dv2 = 3.0f * gamma_b2 * dmb;
    }
    s_b[k] = dv0;
    s_b[kGW * kGW + k] = dv1;
    s_b[2 * kGW * kGW + k] = dv2;
  }
  __syncthreads();

  // 3. Ring positions outside the padded image take the value of the
  //    pixel they mirror: the transpose of the mirrored read folds the
  //    ring back onto the edge, which for the symmetric 3x3 kernel equals
  //    convolving the mirror-extended dG.
  for (int k = tid; k < kGW * kGW; k += 256) {
    const int i = k / kGW, j = k % kGW;
    const int Y = oy - 1 + i, X = ox - 1 + j;
    if (Y >= 0 && Y < P.py && X >= 0 && X < P.px) continue;
    const int si = Mirror(Y, P.py) - oy + 1, sj = Mirror(X, P.px) - ox + 1;
    const bool ok = si >= 0 && si < kGW && sj >= 0 && sj < kGW;
    for (int c = 0; c < 3; ++c) {
      s_b[c * kGW * kGW + k] = ok ? s_b[c * kGW * kGW + si * kGW + sj] : 0.0f;
    }
  }
  __syncthreads();

  // 4. dP = Gaborish transpose (the same symmetric kernel) of the ring.
  for (int k = tid; k < 1024; k += 256) {
    const int i = k >> 5, j = k & 31;
    for (int c = 0; c < 3; ++c) {
      const float* g = s_b + c * kGW * kGW;
      s_pix[c][PixAt(i, j)] = P.gaborish ? Symmetric3(g, kGW, kGW, kGW, j + 1, i + 1,
                                            P.gab_w[0], P.gab_w[1], P.gab_w[2])
                               : g[(i + 1) * kGW + j + 1];
    }
  }
  __syncthreads();

  // 5. Transform transposes: forward transform, then the per-coefficient
  //    norm scale (IDENTITY directly).
  float* s_coef = s_b;
  TileForward3<true>(s_fwd, &s_pix[0][0], s_coef, s_a);

  // 6. CfL transpose, dL/dd and dL/dqf.
  const int gxi = tbx / kGroupDimBlocks, gyi = tby / kGroupDimBlocks;
  const CoefSlot S = MakeSlot(P, s_acs, tbx, tby, gxi * kGroupDimBlocks,
                              gyi * kGroupDimBlocks);
  float part = 0.0f;
  float qfr = 1.0f;
  if (S.valid) {
    const size_t plane = size_t(P.by) * P.bx * 64;
    const size_t at =
        size_t(S.cby) * P.bx * 64 + size_t(S.cbx) * 64 + (S.ccol & 63);
    qfr = A.qf[size_t(S.cby) * A.qf_pitch + S.cbx];
    const float scaled = P.inv_global_scale / qfr;
    const BlockDq dx = MakeBlockDq(P, B, 0, S.cby, S.cbx);
    const BlockDq dy = MakeBlockDq(P, B, 1, S.cby, S.cbx);
    const BlockDq db = MakeBlockDq(P, B, 2, S.cby, S.cbx);
    float out[3][4];
    for (int j = 0; j < 4; ++j) {
      const int flat = S.cflat0 + j;
      const int fi = flat >> S.lgn, fj = flat & (S.cN - 1);
      const int e = CoefAt(S.crow, S.ccol + j);
      if (fi < S.cs && fj < S.cs) {
        out[0][j] = out[1][j] = out[2][j] = 0.0f;
        continue;
      }
      const float sc = AdjointScale(K->adj_norm, S.raw, flat);
      const float gx = s_coef[e] * sc;
      const float gy = s_coef[kCoefPlane + e] * sc;
      const float gb = s_coef[2 * kCoefPlane + e] * sc;
      const float gyt = gy + dx.cc * gx + db.cc * gb;
      const int k = (S.ccol & 63) + j;
      const float vx = DequantPlain(dx.m, scaled, 0, B.ac[at + j], k);
      const float vy = DequantPlain(dy.m, scaled, 1, B.ac[plane + at + j], k);
      const float vb =
          DequantPlain(db.m, scaled, 2, B.ac[2 * plane + at + j], k);
      part += gx * vx + gyt * vy + gb * vb;
      out[0][j] = gx;
      out[1][j] = gyt;
      out[2][j] = gb;
    }
    if (A.dcoef[0]) {
      for (int c = 0; c < 3; ++c) {
        float* dst = A.dcoef[c] + size_t(S.cby) * A.dc_pitch +
                     size_t(S.cbx) * 64 + (S.ccol & 63);
        for (int j = 0; j < 4; ++j) dst[j] = out[c][j];
      }
    }
  }
  // Sum over the 16 lanes of the block (all lanes take part).
  for (int m = 8; m >= 1; m >>= 1) part += __shfl_xor(part, m, 64);
  if (S.valid && (tid & 15) == 0) {
    A.dqf[size_t(S.cby) * A.dqf_pitch + S.cbx] = -part / qfr;
  }

  // 7. Tile loss partial.
  if (A.loss_part) {
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m, 64);
    if ((tid & 63) == 0) s_loss[tid >> 6] = loss;
    __syncthreads();
    if (tid == 0) {
      A.loss_part[tile] = (s_loss[0] + s_loss[1]) + (s_loss[2] + s_loss[3]);
    }
  }
}

// Sum of the per-tile loss partials in double (one workgroup).
__global__ void __launch_bounds__(256)
    KAdjLoss(const float* part, int n, double* out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += double(part[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace pikgpu
