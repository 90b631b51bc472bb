// Device helpers and kernel declarations of the reconstruction (gfx950).
//
// All arithmetic that feeds an integer decision later on (the EPF guide
// bytes and sigma) keeps the reference's operation order with separate
// multiplies and adds; the library is compiled with -ffp-contract=off.
#pragma once

#include "device.h"
#include "xform.h"

namespace pikgpu {

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

// Per-storage-block dequantisation parameters of plane c: the dequant
// matrix row of the block's strategy, position and table, the scaled
// 1 / quant field and the colour-correlation factor
// (Dequant::DoAC, pik/compressed_image.cc:636-666).
struct BlockDq {
  const float* m;
  float scaled;
  float cc;
};

__device__ inline BlockDq MakeBlockDq(const FrameParams& P,
                                      const FrameBuffers& B, int c, int row,
                                      int bxa) {
  const int blk = row * P.bx + bxa;
  const uint8_t a = B.acs[blk];
  const int raw = a >> 4, idx = a & 15;
  const int qf = B.qf[blk];
  const int tile = (row >> 3) * P.tx + (bxa >> 3);
  const int qt = B.dqmap[B.dcf[tile] * 256 + qf - 1];
  BlockDq d;
  d.m = B.table + B.offsets[(qt * kNumQuantKinds + AcsQuantKind(raw)) * 3 + c] +
        idx * 64;
  d.scaled = qf == 0 ? 1E10f : P.inv_global_scale / float(qf);
  d.cc = c == 0 ? float(B.ytox[tile] - 128) * 0.00390625f
                : (c == 2 ? float(B.ytob[tile]) * 0.0078125f : 0.0f);
  return d;
}

// MakeBlockDq's values from the block's descriptor (KDcDequant).
__device__ inline BlockDq DescBlockDq(const FrameParams& P,
                                      const FrameBuffers& B, int c, int row,
                                      int bxa) {
  const uint4 d = B.desc[2 * (row * P.bx + bxa)];
  const int tile = (row >> 3) * P.tx + (bxa >> 3);
  BlockDq q;
  q.m = B.table + (c == 0 ? d.x : (c == 1 ? d.y : d.z));
  q.scaled = __uint_as_float(d.w);
  q.cc = c == 0 ? float(B.ytox[tile] - 128) * 0.00390625f
                : (c == 2 ? float(B.ytob[tile]) * 0.0078125f : 0.0f);
  return q;
}

// Coefficient k of a block from its quantized value q, in DequantCoef's
// operation order.
__device__ inline float DequantWith(const BlockDq& d, int c, int16_t q, int k,
                                    float yval) {
  const float mul = d.m[k] * d.scaled;
  const float v = AdjustQuantBias(c, float(q)) * mul;
  return c == 1 ? v : d.cc * yval + v;
}

// AC plane c of pass p (0: the frame's own, then pikgpu_frame_add_pass's).
__device__ inline const int16_t* AcPlane(const FrameParams& P,
                                         const FrameBuffers& B, int p, int c) {
  const size_t plane = size_t(P.by) * P.bx * 64;
  return p == 0 ? B.ac + c * plane
                : B.ac_extra + (size_t(p - 1) * 3 + c) * plane;
}

// DequantWith summed over the frame's passes: the first pass's value, then
// each later pass's own value (its colour correlation from its own Y) plus
// the sum so far (Dequant<false>::DoAC, pik/compressed_image.cc:695-703).
// `at` indexes the coefficient within a plane; dy is read only for c != 1.
__device__ inline float DequantPasses(const FrameParams& P,
                                      const FrameBuffers& B, const BlockDq& d,
                                      const BlockDq& dy, int c, size_t at,
                                      int k) {
  float acc = 0.0f;
  for (int p = 0; p < P.num_passes; ++p) {
    float yv = 0.0f;
    if (c != 1) yv = DequantWith(dy, 1, AcPlane(P, B, p, 1)[at], k, 0.0f);
    const float v = DequantWith(d, c, AcPlane(P, B, p, c)[at], k, yv);
    acc = p == 0 ? v : v + acc;
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Cooperative transforms of the three channels of a 32x32 tile (256
// threads). The 2-D transforms are the reference's two column passes with a
// transpose in between (pik/dct.h:370-507); each line is a 1-D butterfly of
// dct1d.h computed by one thread. Every line of every transform of every
// channel is a task, and the tasks are spread over all 256 threads: a DCT8
// block has 8 line tasks, DCT16 16, DCT32 32, DCT4X4 16 (4 lines in each
// 4x4 sub-block), DCT2X2 and IDENTITY one direct task.
//
// LDS layouts (per channel):
//  * pixels and the transpose scratch: 32 rows of kTilePitch = 33 floats;
//  * coefficients: the reference's storage order (4 block rows; block bx
//    of a row holds its 64 coefficients, a DCT16/32's F matrix runs across
//    the s blocks of its s rows), with kCoefBlock = 72 floats per block.
// With power-of-two pitches the lines that neighbouring lanes read or write
// would all sit in the same LDS bank (e.g. a DCT8 column pass writing 8
// transposed rows: an 8- to 32-way conflict); the odd pixel pitch and the
// 8-float gap between blocks give the lanes of a wave distinct banks for
// the DCT8/16/32 passes.
constexpr int kTilePitch = 33;
constexpr int kTilePlane = 32 * kTilePitch;  // 1056
constexpr int kCoefBlock = 72;
constexpr int kCoefRow = 4 * kCoefBlock;     // 288
constexpr int kCoefPlane = 4 * kCoefRow;     // 1152

// Pixel (y, x) of a tile channel.
__host__ __device__ inline int PixAt(int y, int x) {
  return y * kTilePitch + x;
}
// Coefficient at storage row `row` (0..3), column `col` (block * 64 + k).
__host__ __device__ inline int CoefAt(int row, int col) {
  return row * kCoefRow + (col >> 6) * kCoefBlock + (col & 63);
}

// The tile's transforms in task order: grouped by kind (DCT8, DCT16,
// DCT32, DCT4X4, direct), so that the lanes of a wave mostly run the same
// line code. Position p holds the transform starting at tile block blk[p].
struct TileTasks {
  int off[17];      // first task of each position; off[16] = tasks per
                    // channel (positions past the last transform: off = total)
  uint8_t raw[16];  // strategy at each position (0xFF past the last)
  uint8_t blk[16];  // tile block (by * 4 + bx) at each position
  int has_dct4;     // a DCT4X4 transform is among them
};

__device__ inline int TaskCount(int raw) {
  if (raw == kDCT || raw == kDCTNoHF) return 8;
  if (raw == kDCT16) return 16;
  if (raw == kDCT32) return 32;
  if (AcsIsDCT4(raw)) return 16;
  return 1;
}

__device__ inline int TaskKind(int raw) {
  if (raw == kDCT || raw == kDCTNoHF) return 0;
  if (raw == kDCT16) return 1;
  if (raw == kDCT32) return 2;
  if (AcsIsDCT4(raw)) return 3;
  return 4;
}

// All threads call; `acs_t` holds the tile's 16 strategy bytes (0xFF
// outside the image). hf_only keeps only strategies with HF prediction.
// Lanes 0..15 of the first wave each place one block: its position is the
// number of transforms of earlier kinds plus those of its kind at earlier
// blocks (ballots), its first task the task count of the positions before
// it (shuffles).
__device__ inline void BuildTasks(const uint8_t* acs_t, bool hf_only,
                                  TileTasks* T) {
  if (threadIdx.x < 64) {
    const int b = threadIdx.x;
    int raw = -1, kind = 5, cnt = 0;
    if (b < 16) {
      const uint8_t a = acs_t[b];
      if (a != 0xFF && (a & 15) == 0 && (!hf_only || AcsPredictHF(a >> 4))) {
        raw = a >> 4;
        kind = TaskKind(raw);
        cnt = TaskCount(raw);
      }
    }
    int pos = 0, n = 0;
    for (int k = 0; k < 5; ++k) {
      const unsigned long long m = __ballot(kind == k);
      n += __popcll(m);
      if (k < kind) pos += __popcll(m);
      if (k == kind) pos += __popcll(m & ((1ull << b) - 1));
    }
    int first = 0, total = 0;
    for (int q = 0; q < 16; ++q) {
      const int qc = __shfl(cnt, q, 64), qp = __shfl(pos, q, 64);
      const int qk = __shfl(kind, q, 64);
      total += qc;
      if (qk < 5 && qp < pos) first += qc;
    }
    if (b < 16 && raw >= 0) {
      T->off[pos] = first;
      T->raw[pos] = uint8_t(raw);
      T->blk[pos] = uint8_t(b);
    }
    if (b < 16 && b >= n) {  // positions past the last transform
      T->off[b] = total;
      T->raw[b] = 0xFF;
      T->blk[b] = 0;
    }
    const unsigned long long m4 = __ballot(kind == 3);
    if (b == 0) {
      T->off[16] = total;
      T->has_dct4 = m4 != 0;
    }
  }
  __syncthreads();
}

// Position of task t (binary search over the offsets).
__device__ inline int TaskBlock(const TileTasks& T, int t) {
  int lo = 0;
  for (int step = 8; step > 0; step >>= 1) {
    if (T.off[lo + step] <= t) lo += step;
  }
  return lo;
}

// Column k of an N x N transform's F matrix (row-major, flat i * N + k)
// with origin block (oby, obx): F[i][k] sits in storage row oby + i / 8,
// column obx * 64 + (i % 8) * N + k, so its index is
//   CoefColumn(oby, obx, k) + CoefRowOffset<N>(i)
// (k < N keeps (i % 8) * N % 64 + k inside one block). With the row offset
// a compile-time constant every access of a line shares one address and
// differs only in the instruction's immediate offset.
__device__ inline unsigned CoefColumn(unsigned oby, unsigned obx, unsigned k) {
  return oby * kCoefRow + obx * kCoefBlock + k;
}
template <int N>
__host__ __device__ constexpr int CoefRowOffset(int i) {
  return (i / 8) * kCoefRow + ((i % 8) * N / 64) * kCoefBlock +
         (i % 8) * N % 64;
}

// CoefRowOffset<N>(i) + k equals the storage index of flat i * N + k
// relative to the origin block for every i, k < N.
template <int N>
constexpr bool CoefRowOffsetsMatchLayout() {
  for (int i = 0; i < N; ++i) {
    for (int k = 0; k < N; ++k) {
      const int flat = i * N + k;
      const int row = flat / (8 * N), col = flat % (8 * N);
      const int idx = row * kCoefRow + (col >> 6) * kCoefBlock + (col & 63);
      if (CoefRowOffset<N>(i) + k != idx) return false;
    }
  }
  return true;
}
static_assert(CoefRowOffsetsMatchLayout<8>() &&
                  CoefRowOffsetsMatchLayout<16>() &&
                  CoefRowOffsetsMatchLayout<32>(),
              "coefficient row offsets");

template <int N>
__device__ inline void InvLine1(const float* coef, float* tmp, int oby,
                                int obx, int c) {
  const float* col = coef + CoefColumn(oby, obx, c);
  float v[N], w[N];
  for (int r = 0; r < N; ++r) v[r] = col[CoefRowOffset<N>(r)];
  dct1d::Idct<N>(v, w);
  for (int x = 0; x < N; ++x) tmp[PixAt(oby * 8 + c, obx * 8 + x)] = w[x];
}
template <int N>
__device__ inline void InvLine2(const float* tmp, float* pix, int oby,
                                int obx, int x) {
  float v[N], w[N];
  for (int c = 0; c < N; ++c) v[c] = tmp[PixAt(oby * 8 + c, obx * 8 + x)];
  dct1d::Idct<N>(v, w);
  for (int y = 0; y < N; ++y) pix[PixAt(oby * 8 + y, obx * 8 + x)] = w[y];
}
template <int N>
__device__ inline void FwdLine1(const float* pix, float* tmp, int oby,
                                int obx, int x) {
  float v[N], w[N];
  for (int y = 0; y < N; ++y) v[y] = pix[PixAt(oby * 8 + y, obx * 8 + x)];
  dct1d::Dct<N>(v, w);
  for (int k = 0; k < N; ++k) tmp[PixAt(oby * 8 + x, obx * 8 + k)] = w[k];
}
// kAddHF: add the coefficients outside the 2s x 2s low-frequency corner
// to coef instead of storing them, or with kSubHF subtract them
// (AddBlockExceptLFAndLLFTo<true / false>, pik/ac_predictions.cc:38-70).
template <int N, bool kAddHF = false, bool kSubHF = false>
__device__ inline void FwdLine2(const float* tmp, float* coef, int oby,
                                int obx, int k) {
  constexpr int s = N / 8;
  float v[N], w[N];
  for (int x = 0; x < N; ++x) v[x] = tmp[PixAt(oby * 8 + x, obx * 8 + k)];
  dct1d::Dct<N>(v, w);
  const float scale = 1.0f / (N * N);
  float* col = coef + CoefColumn(oby, obx, k);
  for (int j = 0; j < N; ++j) {
    float& dst = col[CoefRowOffset<N>(j)];
    if (!kAddHF) {
      dst = w[j] * scale;
    } else if (!(j < 2 * s && k < 2 * s)) {
      dst = kSubHF ? dst - w[j] * scale : dst + w[j] * scale;
    }
  }
}

// One task of the inverse transform: pass 0 columns of the coefficient
// matrix into tmp, pass 1 columns of tmp into pixels (DCT4X4 per 4x4
// sub-block with the Hadamard of the four DCs, pik/ac_strategy.cc:762-790;
// DCT2X2 and IDENTITY directly in pass 1).
__device__ inline void InverseTask(int pass, int raw, int l, int oby, int obx,
                                   const float* coef, float* pix, float* tmp) {
  if (raw == kDCT || raw == kDCTNoHF) {
    if (pass == 0) InvLine1<8>(coef, tmp, oby, obx, l);
    else InvLine2<8>(tmp, pix, oby, obx, l);
  } else if (raw == kDCT16) {
    if (pass == 0) InvLine1<16>(coef, tmp, oby, obx, l);
    else InvLine2<16>(tmp, pix, oby, obx, l);
  } else if (raw == kDCT32) {
    if (pass == 0) InvLine1<32>(coef, tmp, oby, obx, l);
    else InvLine2<32>(tmp, pix, oby, obx, l);
  } else if (AcsIsDCT4(raw)) {
    const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, q = l & 3;
    float v[4], w[4];
    if (pass == 0) {
      // column q of the sub-block's 4x4 matrix G, G[iy][ix] =
      // coef[(sy + 2 iy) * 8 + sx + 2 ix], G[0][0] from the DC Hadamard
      const float* cb = coef + CoefAt(oby, obx * 64);
      const float b00 = cb[0], b01 = cb[1], b10 = cb[8], b11 = cb[9];
      float dc;
      if (sub == 0) dc = b00 + b01 + b10 + b11;
      else if (sub == 1) dc = b00 + b01 - b10 - b11;
      else if (sub == 2) dc = b00 - b01 + b10 - b11;
      else dc = b00 - b01 - b10 + b11;
      for (int r = 0; r < 4; ++r) {
        v[r] = (r == 0 && q == 0) ? dc : cb[(sy + r * 2) * 8 + sx + q * 2];
      }
      dct1d::Idct<4>(v, w);
      for (int x = 0; x < 4; ++x) {
        tmp[PixAt(oby * 8 + sy * 4 + q, obx * 8 + sx * 4 + x)] = w[x];
      }
    } else {
      for (int c = 0; c < 4; ++c) {
        v[c] = tmp[PixAt(oby * 8 + sy * 4 + c, obx * 8 + sx * 4 + q)];
      }
      dct1d::Idct<4>(v, w);
      for (int y = 0; y < 4; ++y) {
        pix[PixAt(oby * 8 + sy * 4 + y, obx * 8 + sx * 4 + q)] = w[y];
      }
    }
  } else if (pass == 1) {
    DirectInverse8(raw, coef + CoefAt(oby, obx * 64),
                   pix + PixAt(oby * 8, obx * 8), kTilePitch);
  }
}

// One task of the forward transform: pass 0 columns of pixels into tmp,
// pass 1 columns of tmp into coefficients (x 1/N^2), pass 2 the DC Hadamard
// of DCT4X4 (pik/ac_strategy.cc:590-613); DCT2X2 and IDENTITY in pass 0.
// kAdjoint replaces IDENTITY's forward transform by the transpose of its
// inverse (the other strategies' transposes are scaled forward transforms).
// kAddHF adds the result outside each transform's 2s x 2s low-frequency
// corner to coef (the HF prediction; only strategies with PredictHF, so no
// DCT2X2 or IDENTITY, and the DCT4X4 DC Hadamard, which only touches that
// corner, is skipped).
template <bool kAdjoint = false, bool kAddHF = false, bool kSubHF = false>
__device__ inline void ForwardTask(int pass, int raw, int l, int oby, int obx,
                                   const float* pix, float* coef, float* tmp) {
  if (raw == kDCT || raw == kDCTNoHF) {
    if (pass == 0) FwdLine1<8>(pix, tmp, oby, obx, l);
    else if (pass == 1) FwdLine2<8, kAddHF, kSubHF>(tmp, coef, oby, obx, l);
  } else if (raw == kDCT16) {
    if (pass == 0) FwdLine1<16>(pix, tmp, oby, obx, l);
    else if (pass == 1) FwdLine2<16, kAddHF, kSubHF>(tmp, coef, oby, obx, l);
  } else if (raw == kDCT32) {
    if (pass == 0) FwdLine1<32>(pix, tmp, oby, obx, l);
    else if (pass == 1) FwdLine2<32, kAddHF, kSubHF>(tmp, coef, oby, obx, l);
  } else if (AcsIsDCT4(raw)) {
    const int sub = l >> 2, sy = sub >> 1, sx = sub & 1, q = l & 3;
    float v[4], w[4];
    float* cb = coef + CoefAt(oby, obx * 64);
    if (pass == 0) {
      for (int y = 0; y < 4; ++y) {
        v[y] = pix[PixAt(oby * 8 + sy * 4 + y, obx * 8 + sx * 4 + q)];
      }
      dct1d::Dct<4>(v, w);
      for (int k = 0; k < 4; ++k) {
        tmp[PixAt(oby * 8 + sy * 4 + q, obx * 8 + sx * 4 + k)] = w[k];
      }
    } else if (pass == 1) {
      for (int x = 0; x < 4; ++x) {
        v[x] = tmp[PixAt(oby * 8 + sy * 4 + x, obx * 8 + sx * 4 + q)];
      }
      dct1d::Dct<4>(v, w);
      for (int j = 0; j < 4; ++j) {
        float& dst = cb[(sy + j * 2) * 8 + sx + q * 2];
        if (!kAddHF) {
          dst = w[j] * (1.0f / 16);
        } else if (!(j == 0 && q == 0)) {  // outside the 2x2 corner
          dst = kSubHF ? dst - w[j] * (1.0f / 16) : dst + w[j] * (1.0f / 16);
        }
      }
    } else if (!kAddHF && l == 0) {
      const float b00 = cb[0], b01 = cb[1], b10 = cb[8], b11 = cb[9];
      cb[0] = (b00 + b01 + b10 + b11) * 0.25f;
      cb[1] = (b00 + b01 - b10 - b11) * 0.25f;
      cb[8] = (b00 - b01 + b10 - b11) * 0.25f;
      cb[9] = (b00 - b01 - b10 + b11) * 0.25f;
    }
  } else if (pass == 0) {
    if (kAdjoint && raw == kIdentity) {
      IdentityAdjoint8(pix + PixAt(oby * 8, obx * 8), kTilePitch,
                       coef + CoefAt(oby, obx * 64));
    } else {
      DirectForward8(raw, pix + PixAt(oby * 8, obx * 8), kTilePitch,
                     coef + CoefAt(oby, obx * 64));
    }
  }
}

// A thread's share of a tile's tasks, decoded once for all passes: task
// k = threadIdx.x + 256 * i (channel-major over the 3 channels) in slot i,
// packed as channel | raw << 2 | line << 6 | oby << 11 | obx << 13, and
// kNoTask past the last task. A channel has at most 16 blocks x 16 line
// tasks (DCT4X4), so three slots cover 3 x 256 tasks.
constexpr int kTaskSlots = 3;
constexpr uint32_t kNoTask = 0xFFFFFFFFu;
struct ThreadTasks {
  uint32_t t[kTaskSlots];
  bool has_dct4;  // TileTasks::has_dct4 (the same for the workgroup)
  // slot i by selects (a variable index into t would put it in scratch)
  __device__ uint32_t Slot(int i) const {
    return i == 0 ? t[0] : (i == 1 ? t[1] : t[2]);
  }
};

__device__ inline ThreadTasks AssignTasks(const TileTasks& T) {
  const int total = T.off[16];
  ThreadTasks m;
  m.has_dct4 = T.has_dct4 != 0;
#pragma unroll
  for (int i = 0; i < kTaskSlots; ++i) {
    const int k = threadIdx.x + 256 * i;
    m.t[i] = kNoTask;
    if (k < 3 * total) {
      const int c = (k >= total) + (k >= 2 * total), t = k - c * total;
      const int p = TaskBlock(T, t), b = T.blk[p];
      m.t[i] = uint32_t(c) | (uint32_t(T.raw[p]) << 2) |
               (uint32_t(t - T.off[p]) << 6) | (uint32_t(b >> 2) << 11) |
               (uint32_t(b & 3) << 13);
    }
  }
  return m;
}

// coef (3 x kCoefPlane) -> pix (3 x kTilePlane), tmp 3 x kTilePlane. All
// 256 threads must call.
__device__ inline void TileInverse3(const ThreadTasks& m, const float* coef,
                                    float* pix, float* tmp) {
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll 1
    for (int i = 0; i < kTaskSlots; ++i) {
      const uint32_t w = m.Slot(i);
      if (w != kNoTask) {
        const int c = w & 3;
        InverseTask(pass, (w >> 2) & 15, (w >> 6) & 31, (w >> 11) & 3,
                    (w >> 13) & 3, coef + c * kCoefPlane, pix + c * kTilePlane,
                    tmp + c * kTilePlane);
      }
    }
    __syncthreads();
  }
}
__device__ inline void TileInverse3(const TileTasks& T, const float* coef,
                                    float* pix, float* tmp) {
  TileInverse3(AssignTasks(T), coef, pix, tmp);
}

// pix (3 x kTilePlane) -> coef (3 x kCoefPlane) for the transforms in T
// (others keep their coefficient entries), tmp 3 x kTilePlane. All 256
// threads must call.
template <bool kAdjoint = false, bool kAddHF = false, bool kSubHF = false>
__device__ inline void TileForward3(const ThreadTasks& m, const float* pix,
                                    float* coef, float* tmp) {
  // pass 2 (the DCT4X4 DC Hadamard) only when the tile has a DCT4X4
  const int passes = kAddHF ? 2 : (m.has_dct4 ? 3 : 2);
  for (int pass = 0; pass < passes; ++pass) {
#pragma unroll 1
    for (int i = 0; i < kTaskSlots; ++i) {
      const uint32_t w = m.Slot(i);
      if (w != kNoTask) {
        const int c = w & 3;
        ForwardTask<kAdjoint, kAddHF, kSubHF>(
            pass, (w >> 2) & 15, (w >> 6) & 31, (w >> 11) & 3,
            (w >> 13) & 3, pix + c * kTilePlane, coef + c * kCoefPlane,
            tmp + c * kTilePlane);
      }
    }
    __syncthreads();
  }
}
template <bool kAdjoint = false, bool kAddHF = false, bool kSubHF = false>
__device__ inline void TileForward3(const TileTasks& T, const float* pix,
                                    float* coef, float* tmp) {
  TileForward3<kAdjoint, kAddHF, kSubHF>(AssignTasks(T), pix, coef, tmp);
}

// ---------------------------------------------------------------------------
// Workgroup-to-tile order for the eight XCDs of an MI355X. Consecutive
// workgroup ids are dealt round-robin to the XCDs, each with its own L2, so
// in plain row-major order neighbouring tiles -- which read each other's
// halo rows and columns -- sit in different L2s. Here XCD k walks the
// contiguous row-major tile range [k * per, (k + 1) * per) in order.
// Launch 8 * per workgroups (per = ceil(num_tiles / 8)); ids past the last
// tile get -1 and return at once.
constexpr int kNumXcds = 8;

__host__ __device__ inline int XcdGrid(int num_tiles) {
  return kNumXcds * ((num_tiles + kNumXcds - 1) / kNumXcds);
}

__device__ inline int XcdTile(int wg, int num_tiles) {
  const int per = (num_tiles + kNumXcds - 1) / kNumXcds;
  const int tile = (wg % kNumXcds) * per + wg / kNumXcds;
  return tile < num_tiles ? tile : -1;
}

// ---------------------------------------------------------------------------
// The workgroup a kernel body works on: blockIdx / gridDim of a plain
// launch, or, in a batched launch, the frame's own grid inside one launch
// over every frame of the batch (BatchSpan).
struct VBlock {
  int x, y, z;     // workgroup index
  int nx, ny, nz;  // grid size in workgroups
};
__device__ inline VBlock PlainBlock() {
  VBlock v;
  v.x = int(blockIdx.x);
  v.y = int(blockIdx.y);
  v.z = int(blockIdx.z);
  v.nx = int(gridDim.x);
  v.ny = int(gridDim.y);
  v.nz = int(gridDim.z);
  return v;
}

// A frame of a batch in device memory: its parameters and buffers as a
// plain launch passes them.
struct BatchFrame {
  FrameParams P;
  FrameBuffers B;
};
// One frame's share of a batched launch of one phase: launch workgroups
// [first, first + count) run the frame's nx x ny x nz grid (count =
// nx * ny * nz), with the phase's per-frame integer arguments. `first` is
// a multiple of kNumXcds, so a frame's local workgroup index keeps the XCD
// of the launch's (XcdTile).
constexpr int kBatchArgs = 5;
struct BatchSpan {
  uint32_t frame, first, count;
  int nx, ny, nz;
  int arg[kBatchArgs];
};

// The span table's companion map: entry w is the index (into the launch's
// spans) of the span holding launch workgroup w, or kNoSpan for the idle
// workgroups between spans (one load instead of a search per workgroup).
typedef const BatchSpan ConstSpan;
constexpr uint32_t kNoSpan = 0xFFFFFFFFu;

__device__ inline VBlock SpanBlock(ConstSpan& s, uint32_t w) {
  const int l = int(w - s.first);
  VBlock v;
  v.nx = s.nx;
  v.ny = s.ny;
  v.nz = s.nz;
  v.x = l % s.nx;
  v.y = (l / s.nx) % s.ny;
  v.z = l / (s.nx * s.ny);
  return v;
}

// The loop of a batched kernel: workgroup blockIdx.x takes launch
// workgroups blockIdx.x, + gridDim.x, ... below `total` (one of them when
// the launch has `total` workgroups: a persistent launch has fewer), runs
// `body` with s (the BatchSpan), vb (the frame's workgroup), P and B (the
// frame's), and waits for the workgroup before the next one reuses LDS.
#define PIKGPU_BATCH_LOOP(body)                                       \
  for (uint32_t w_ = blockIdx.x; w_ < total; w_ += gridDim.x) {       \
    const uint32_t si_ = M[w_];                                       \
    if (si_ != kNoSpan) {                                             \
      ConstSpan& s = S[si_];                                          \
      const VBlock vb = SpanBlock(s, w_);                             \
      const FrameParams& P = F[s.frame].P;                            \
      const FrameBuffers& B = F[s.frame].B;                           \
      body;                                                           \
    }                                                                 \
    __syncthreads();                                                  \
  }

// One launch workgroup per frame workgroup (a launch of `total`
// workgroups), for the kernels whose registers a loop around them would
// overflow (the compiler keeps per-thread values of the body live across
// the iterations).
#define PIKGPU_BATCH_ONE(body)                                        \
  {                                                                   \
    const uint32_t w_ = blockIdx.x;                                   \
    const uint32_t si_ = w_ < total ? M[w_] : kNoSpan;                \
    if (si_ != kNoSpan) {                                             \
      ConstSpan& s = S[si_];                                          \
      const VBlock vb = SpanBlock(s, w_);                             \
      const FrameParams& P = F[s.frame].P;                            \
      const FrameBuffers& B = F[s.frame].B;                           \
      body;                                                           \
    }                                                                 \
  }

// The batched kernels' common parameters.
#define PIKGPU_BATCH_PARAMS                                             \
  const BatchFrame *__restrict__ F, const BatchSpan *__restrict__ S,    \
      const uint32_t *__restrict__ M, uint32_t total,                   \
      const DeviceConsts *K


// ---------------------------------------------------------------------------
// Small reinterpreting transforms with compile-time sizes: every array index
// is static after unrolling, so the blocks stay in registers.

// 2-point transforms in the reference's explicit form (pik/dct.h:478-504).
__device__ inline void Idct2x2(float a00, float a01, float a10, float a11,
                               float* o, int os) {
  o[0] = a00 + a01 + a10 + a11;
  o[1] = a00 + a01 - a10 - a11;
  o[os] = a00 - a01 + a10 - a11;
  o[os + 1] = a00 - a01 - a10 + a11;
}

// Index of the scale tables for size n: log2(n) - 1.
template <int N>
constexpr int ScaleIdx() {
  return N == 2 ? 0 : (N == 4 ? 1 : (N == 8 ? 2 : (N == 16 ? 3 : 4)));
}

// ReinterpretingDCT<DCTN, N, N> (pik/ac_strategy.cc:272-293), N >= 4:
// N x N input at `in` (pitch ip) -> N x N output at `out` (pitch op); in and
// out may alias.
template <int DCTN, int N>
__device__ inline void ReinterpDct(const DeviceConsts* K, const float* in,
                                   int ip, float* out, int op) {
  float blk[N * N], coef[N * N];
  for (int y = 0; y < N; ++y) {
    for (int x = 0; x < N; ++x) blk[y * N + x] = in[y * ip + x];
  }
  dct1d::Dct2D<N>(blk, N, coef, N);
  constexpr int ti = ScaleIdx<N>(), ai = ScaleIdx<DCTN>();
  for (int y = 0; y < N; ++y) {
    for (int x = 0; x < N; ++x) {
      const float v = coef[y * N + x] * K->tscale[ti][y][x];
      out[y * op + x] = v * K->ascale[ai][y][x];
    }
  }
}

// ReinterpretingIDCT<DCTN, LF, N> (pik/ac_strategy.cc:296-315), N >= 4:
// the LF x LF input (pitch ip) zero-padded to N x N.
template <int DCTN, int LF, int N>
__device__ inline void ReinterpIdct(const DeviceConsts* K, const float* in,
                                    int ip, float* out, int op) {
  float blk[N * N], pix[N * N];
  constexpr int ai = ScaleIdx<N>(), ti = ScaleIdx<DCTN>();
#pragma unroll
  for (int y = 0; y < N; ++y) {
#pragma unroll
    for (int x = 0; x < N; ++x) {
      float v = 0.0f;
      if (y < LF && x < LF) {
        v = (in[y * ip + x] * K->ascale[ai][y][x]) * K->tscale[ti][y][x];
      }
      blk[y * N + x] = v;
    }
  }
  dct1d::Idct2D<N>(blk, N, pix, N);
#pragma unroll
  for (int y = 0; y < N; ++y) {
#pragma unroll
    for (int x = 0; x < N; ++x) out[y * op + x] = pix[y * N + x];
  }
}

// Quantizer::QuantizeDC's factor per channel: InvMatrix(0, DCT8, c)[0] *
// (Scale() * quant_dc) (pik/quantizer.h:350-353).
struct EncDcScale {
  float scale[3];
};
// KEncQuant's workgroup size (its LDS holds 64 floats per thread).
constexpr int kEncQuantThreads = 128;

// ---------------------------------------------------------------------------
// Kernels.
__global__ void KDcDequant(FrameParams P, FrameBuffers B);
__global__ void KEncForward(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                            const float* opsin, float* coef, float* dcf);
__global__ void KEncDc(FrameParams P, const float* dcf, EncDcScale S,
                       int16_t* dc16);
__global__ void KEncQuant(FrameParams P, FrameBuffers B, const float* coef,
                          const float* itable, float gsf, int16_t* ac);
__global__ void KEncLf(FrameParams P, FrameBuffers B, float* coef,
                       const float* itable, float gsf, float* corner);
__global__ void KEncHfResidual(FrameParams P, FrameBuffers B, float* coef);
__global__ void KComputeLlf(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                            int dc2x2, int wg4, int wg2, int pos_blocks,
                            int ngroups);
__global__ void KDc2x2Border(FrameParams P, FrameBuffers B,
                             const DeviceConsts* K);
__global__ void KLfBlur(FrameParams P, FrameBuffers B);
__global__ void KLfFromDc2x2(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                             int wg4, int wg2, int pos_blocks, int ngroups);
template <bool kUseLf>
__global__ void KPred2x2Border(FrameParams P, FrameBuffers B, const DeviceConsts* K);
template <bool kEnc>
__global__ void KPred2x2Interior(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                                 int wg4, int wg2, const float* corner);
__global__ void KBlocks(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KAdjForward(FrameParams P, FrameBuffers B,
                            const DeviceConsts* K, AdjArgs A);
__global__ void KAdjBackward(FrameParams P, FrameBuffers B,
                             const DeviceConsts* K, AdjArgs A);
__global__ void KAdjLoss(const float* part, int n, double* out);

// Parameters of one Adam step of the quant-field search (search.hip).
struct SearchArgs {
  const float* weight;  // per-block perceptual weight w_b, or null (1)
  const float* lambda;  // rate weight per block (device scalar)
  float lr, beta1, beta2, eps;
  float ulo, uhi;       // bounds of the log field
  float dlo, dhi;       // trust region around the initial log field
};
__global__ void KSearchRound(FrameParams P, FrameBuffers B, const float* u,
                             int qmax, int32_t* qf, float* qfr, float* part);
__global__ void KSearchDist(FrameParams P, FrameBuffers B, const float* coef,
                            const float* grad, float* dist);
__global__ void KSearchUpdate(FrameParams P, FrameBuffers B, SearchArgs S,
                              int t, const float* dist, const float* u0,
                              float* u, float* m, float* v);
__global__ void KSearchInit(FrameParams P, FrameBuffers B, float* u0, float* u,
                            float* m, float* v);
__global__ void KSearchLambda(const float* dist, const float* weight, int n,
                              float fixed, float kappa, float* out);

// Generic image description for the filters.
struct Img {
  const float* p;  // 3 planes
  size_t plane;    // floats per plane
  int pitch;       // floats per row
  int W, H;        // logical size (mirror borders)
  int vw, vh;      // valid source region (replicate beyond, <= W, H)
};

// Images of a frame as the filters take them: the IDCT image (full size,
// width a multiple of 8, rows 32-float aligned) and the 8-padded DC image
// (valid bx x by).
__host__ __device__ inline Img IdctImage(const FrameParams& P,
                                         const FrameBuffers& B) {
  Img ii;
  ii.p = B.idct;
  ii.plane = size_t(B.ipitch) * P.py;
  ii.pitch = B.ipitch;
  ii.W = P.px;
  ii.H = P.py;
  ii.vw = P.px;
  ii.vh = P.py;
  return ii;
}
__host__ __device__ inline Img DcImage(const FrameParams& P,
                                       const FrameBuffers& B) {
  Img dci;
  dci.p = B.dcq;
  dci.plane = size_t(P.bx) * P.by;
  dci.pitch = P.bx;
  dci.W = (P.bx + 7) / 8 * 8;
  dci.H = (P.by + 7) / 8 * 8;
  dci.vw = P.bx;
  dci.vh = P.by;
  return dci;
}

__global__ void KMinMaxPartial(Img in, int gaborish, FrameParams P,
                               float* partial);
__global__ void KMinMaxPartial4(Img in, int gaborish, FrameParams P,
                               float* partial);
__global__ void KMinMaxFinal(const float* partial, int n, float* mm);
template <bool kDc>
__global__ void KEpf(Img in, FrameParams P, FrameBuffers B,
                     const DeviceConsts* K, const float* mm, float* out,
                     int out_pitch, size_t out_plane);
__global__ void KGuide(Img in, int gaborish, FrameParams P, FrameBuffers B,
                       const float* mm);
__global__ void KGuideRing(FrameParams P, FrameBuffers B);
__global__ void KDots(FrameParams P, FrameBuffers B);
// KFinal: rows per workgroup of its 256-column strips.
constexpr int kFinalRows = 16;
__global__ void KFinal(FrameParams P, FrameBuffers B, const DeviceConsts* K);
__global__ void KCopyPlanes(const float* src, float* dst, size_t n);

// Batched forms of the reconstruction's kernels (pikgpu_batch): one launch
// per phase over every frame of a batch; see BatchSpan.
__global__ void KDcDequantBatch(PIKGPU_BATCH_PARAMS);
__global__ void KMinMaxPartialDcBatch(PIKGPU_BATCH_PARAMS);
__global__ void KMinMaxFinalBatch(PIKGPU_BATCH_PARAMS);
__global__ void KEpfDcBatch(PIKGPU_BATCH_PARAMS);
__global__ void KComputeLlfBatch(PIKGPU_BATCH_PARAMS);
__global__ void KLfBlurBatch(PIKGPU_BATCH_PARAMS);
__global__ void KLfFromDc2x2Batch(PIKGPU_BATCH_PARAMS);
__global__ void KPred2x2BorderBatch(PIKGPU_BATCH_PARAMS);
__global__ void KDc2x2BorderBatch(PIKGPU_BATCH_PARAMS);
__global__ void KPred2x2InteriorBatch(PIKGPU_BATCH_PARAMS);
__global__ void KBlocksBatch(PIKGPU_BATCH_PARAMS);
__global__ void KMinMaxPartial4Batch(PIKGPU_BATCH_PARAMS);
__global__ void KGuideBatch(PIKGPU_BATCH_PARAMS);
__global__ void KGuideRingBatch(PIKGPU_BATCH_PARAMS);
__global__ void KEpfBatch(PIKGPU_BATCH_PARAMS);
__global__ void KCopyPlanesBatch(PIKGPU_BATCH_PARAMS);
__global__ void KDotsBatch(PIKGPU_BATCH_PARAMS);
__global__ void KFinalBatch(PIKGPU_BATCH_PARAMS);

}  // namespace pikgpu
