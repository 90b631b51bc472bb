// Encoder-side quantization of an XYB image under an uploaded frame's side
// information: the coefficient half of the reference's RoundtripImage
// (pik/adaptive_quantization.cc:580-651), i.e. InitializeFrameEncCache
// (pik/compressed_image.cc:210-304) and ComputeCoefficients (:351-455), for
// frames with and without LF / HF prediction. Its int16 AC and DC land in
// the context's coefficient buffers, so the decode that follows is the round
// trip.
//
//   KEncForward  per 32x32 tile: TransformFromPixels of every transform
//                (the cooperative LDS transforms of the decoder's AR step)
//                and DCFromLowestFrequencies;
//   KEncDc       per block: QuantizeRoundtripDC of Y, the DC colour
//                correlation removed from X and B, QuantizeCoeffsDC;
//   (the decoder's DequantDC + AdaptiveDCReconstruction, ComputeLlf and
//    PredictLf then give dc_dec, the LLF that CopyLlf puts into the
//    coefficients and the LF prediction)
//   KEncLf       per block: LF residuals and the decoder-side LF corner
//                (PredictLfForEncoder), whose 2x DC image KPred2x2Interior
//                forms; KEncHfResidual per tile: the HF residuals
//                (ComputePredictionResiduals);
//   KEncQuant    per block: one 8x8 dither sub-block of its transform, Y
//                quantized and dequantised (QuantizeRoundtripBlockAC), X and
//                B with Y's colour correlation removed
//                (UnapplyColorCorrelationAC), every channel quantized with
//                QuantizeBlockAC's carried rounding error.
#include "device.h"
#include "kernels.h"

namespace pikgpu {

// TransformFromPixels + DCFromLowestFrequencies (pik/compressed_image.cc:
// 234-247, pik/ac_strategy.cc:889-927). `opsin`: 3 planes of py rows x px;
// `coef`: 3 planes of by rows x bx*64; `dcf`: 3 planes of by x bx.
__global__ void __launch_bounds__(256)
    KEncForward(FrameParams P, FrameBuffers B, const DeviceConsts* K,
                const float* opsin, float* coef, float* dcf) {
  __shared__ float s_pix[3][kTilePlane];
  __shared__ float s_tmp[3][kTilePlane];
  __shared__ float s_coef[3][kCoefPlane];
  __shared__ uint8_t s_acs[16];
  __shared__ TileTasks s_tasks;
  const int tiles_x = (P.bx + 3) / 4;
  const int tx4 = blockIdx.x % tiles_x, ty4 = blockIdx.x / tiles_x;
  const size_t plane = size_t(P.px) * P.py;
  const size_t nb = size_t(P.bx) * P.by, row = size_t(P.bx) * 64;
  if (threadIdx.x < 16) {
    const int bxx = tx4 * 4 + (threadIdx.x & 3);
    const int byy = ty4 * 4 + (threadIdx.x >> 2);
    s_acs[threadIdx.x] = (bxx < P.bx && byy < P.by)
                             ? B.acs[size_t(byy) * P.bx + bxx]
                             : uint8_t(0xFF);
  }
  for (int i = threadIdx.x; i < 3 * 32 * 32; i += 256) {
    const int c = i >> 10, r = (i >> 5) & 31, col = i & 31;
    const int y = ty4 * 32 + r, x = tx4 * 32 + col;
    s_pix[c][PixAt(r, col)] =
        (x < P.px && y < P.py) ? opsin[c * plane + size_t(y) * P.px + x] : 0.f;
  }
  __syncthreads();
  BuildTasks(s_acs, false, &s_tasks);
  TileForward3(AssignTasks(s_tasks), &s_pix[0][0], &s_coef[0][0],
               &s_tmp[0][0]);
  for (int i = threadIdx.x; i < 3 * 16 * 64; i += 256) {
    const int c = i >> 10, blk = (i >> 6) & 15, k = i & 63;
    const int r = blk >> 2, b = blk & 3;
    const int y = ty4 * 4 + r, x = tx4 * 4 + b;
    if (x < P.bx && y < P.by) {
      coef[c * nb * 64 + size_t(y) * row + size_t(x) * 64 + k] =
          s_coef[c][CoefAt(r, b * 64 + k)];
    }
  }
  if (threadIdx.x < 48) {
    const int c = threadIdx.x >> 4, blk = threadIdx.x & 15;
    const uint8_t a = s_acs[blk];
    if (a != 0xFF && (a & 15) == 0) {
      const int raw = a >> 4, r = blk >> 2, b = blk & 3;
      const int y = ty4 * 4 + r, x = tx4 * 4 + b;
      const float* F = &s_coef[c][CoefAt(r, b * 64)];
      float* d = dcf + c * nb + size_t(y) * P.bx + x;
      if (raw == kDCT16) {
        // ReinterpretingIDCT<16, 2, 2> of F[0..1][0..1] (row pitch 16)
        constexpr int ti = ScaleIdx<16>();
        const float a00 = (F[0] * K->ascale[0][0][0]) * K->tscale[ti][0][0];
        const float a01 = (F[1] * K->ascale[0][0][1]) * K->tscale[ti][0][1];
        const float a10 = (F[16] * K->ascale[0][1][0]) * K->tscale[ti][1][0];
        const float a11 = (F[17] * K->ascale[0][1][1]) * K->tscale[ti][1][1];
        Idct2x2(a00, a01, a10, a11, d, P.bx);
      } else if (raw == kDCT32) {
        // F[i][k] = storage column i * 32 + k of the first row: rows 2 and
        // 3 lie in the next block's LDS slot (72-float pitch)
        float lo[16], o[16];
        for (int i = 0; i < 4; ++i) {
          for (int k = 0; k < 4; ++k) {
            lo[i * 4 + k] = s_coef[c][CoefAt(r, b * 64 + i * 32 + k)];
          }
        }
        ReinterpIdct<32, 4, 4>(K, lo, 4, o, 4);
        for (int yy = 0; yy < 4; ++yy) {
          for (int xx = 0; xx < 4; ++xx) d[yy * P.bx + xx] = o[yy * 4 + xx];
        }
      } else {
        d[0] = F[0];
      }
    }
  }
}

// QuantizeRoundtripDC of Y, ApplyColorCorrelationDC<false> and
// QuantizeCoeffsDC (pik/compressed_image.cc:262-271, pik/quantizer.cc:
// 86-117, pik/color_correlation.cc:180-205). scale[c] = InvMatrix(0, DCT8,
// c)[0] * (Scale() * quant_dc), the factor of Quantizer::QuantizeDC.
__global__ void __launch_bounds__(256)
    KEncDc(FrameParams P, const float* dcf, EncDcScale S, int16_t* dc16) {
  const size_t nb = size_t(P.bx) * P.by;
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nb) return;
  const int qy = int(roundf(dcf[nb + i] * S.scale[1]));
  const float ydec = qy * P.mul_dc[1];
  const float xv = dcf[i] - P.ytox_dc * ydec;
  const float bv = dcf[2 * nb + i] - P.ytob_dc * ydec;
  dc16[i] = int16_t(int(roundf(xv * S.scale[0])));
  dc16[nb + i] = int16_t(qy);
  dc16[2 * nb + i] = int16_t(int(roundf(bv * S.scale[2])));
}

// AdjustQuantBiasVar (pik/quantizer.h:66-74) of an integer.
__device__ inline float AdjustQuantBiasInt(int c, int q) {
  constexpr float kOneBias[3] = {0.05465007330715401f, 0.07005449891748593f,
                                 0.049935103337343655f};
  if (q == 0) return 0.f;
  if (q == 1) return 1 - kOneBias[c];
  if (q == -1) return kOneBias[c] - 1;
  return q - 0.145f / q;
}

// One step k = 8 * row + j of QuantizeBlockAC's sub-block loop
// (pik/quantizer.h:233-283): the rounding error carried in double along the
// row (err) and from the row above (prev[j]).
template <int j>
__device__ inline int16_t DitherStep(float val, float thres, bool k0,
                                     double& err, double (&prev)[8]) {
  if (j == 0) {
    err = prev[0];
  } else {
    err += prev[j];
  }
  if (err > 0) {
    if (val > 0) {
      val += 0.5 * err;
    } else {
      val -= 0.5 * err;
    }
    err = 0;
  }
  if (fabsf(val) > 1) err = 0;
  double v = (fabsf(val) < thres) ? 0.0 : double(roundf(val));
  if (fabs(v) < fabs(double(val))) {
    err += fabs(v - val);
  } else {
    err -= fabs(v - val);
  }
  if (k0) err = 0;
  if (j == 7) {
    prev[j] = 1.0 * err;
    err = 0;
  } else {
    err *= 0.5;
    prev[j] = err;
  }
  if (v > 32767) v = 32767;
  if (v < -32767) v = -32767;
  return int16_t(v);
}

// Position k of dither sub-block (ix, iy) of an s x s transform whose first
// block is (x0, y0): its index `at` in a coefficient plane (by rows of
// bx*64), its index `pos` in the transform's dequant matrix, and its
// storage row and block column (for the colour-correlation tile).
struct SubPos {
  size_t at;
  int pos, row, col_block;
};
__device__ inline SubPos SubPosition(int k, int s, int ix, int iy, int x0,
                                     int y0, size_t row_len) {
  const int xx = s * (k & 7) + ix, yy = s * (k >> 3) + iy;
  const int pos = yy * 8 * s + xx;
  const int r = yy >> 3, col = pos - r * 64 * s;
  SubPos p;
  p.pos = pos;
  p.row = y0 + r;
  p.col_block = x0 + (col >> 6);
  p.at = size_t(y0 + r) * row_len + size_t(x0) * 64 + col;
  return p;
}

// Y's 64 positions of the sub-block, then X and B with Y's dequantised
// values' colour correlation removed (pik/compressed_image.cc:396-453).
// `coef` holds the transforms' coefficients; the LLF position (k = 0 of
// every sub-block: F[iy][ix]) comes from the group's LLF cache (CopyLlf,
// pik/ac_predictions.cc:581-607). Y's dequantised values wait in LDS, one
// column per thread.
__global__ void __launch_bounds__(kEncQuantThreads)
    KEncQuant(FrameParams P, FrameBuffers B, const float* coef,
              const float* itable, float gsf, int16_t* ac) {
  __shared__ float s_ydec[64][kEncQuantThreads];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kEncQuantThreads + t;
  if (i >= P.bx * P.by) return;
  const int y = i / P.bx, x = i - y * P.bx;
  const uint8_t a = B.acs[i];
  const int raw = a >> 4, idx = a & 15, s = AcsCovered(raw);
  const int iy = idx / s, ix = idx - iy * s;
  const int x0 = x - ix, y0 = y - iy;
  const int quant = B.qf[size_t(y0) * P.bx + x0];
  const int cf = B.dcf[(y0 >> 3) * P.tx + (x0 >> 3)];
  const int table = B.dqmap[cf * 256 + quant];
  const int kind = AcsQuantKind(raw);
  const float qac = gsf * quant;
  const float inv_qac = P.inv_global_scale / quant;
  const size_t nb = size_t(P.bx) * P.by, row_len = size_t(P.bx) * 64;
  const int g = (y0 / kGroupDimBlocks) * P.gx + x0 / kGroupDimBlocks;
  const int lx = x0 % kGroupDimBlocks + 1 + ix,
            ly = y0 % kGroupDimBlocks + 1 + iy;
  constexpr float kZeroBias[3] = {0.65f, 0.6f, 0.7f};

  // channel order Y, X, B: X and B need Y's dequantised values
#pragma unroll 1
  for (int ci = 0; ci < 3; ++ci) {
    const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
    const uint32_t off = B.offsets[(table * kNumQuantKinds + kind) * 3 + c];
    const float* dm = B.table + off;
    const float* qm = itable + off;
    const int32_t* cmap = c == 0 ? B.ytox : B.ytob;
    const int cmap_off = c == 0 ? 128 : 0;
    const float cmap_scale = c == 0 ? 1.0f / 256 : 1.0f / 128;
    const float llf =
        B.llf[(size_t(g) * 3 + c) * kLlfPlane + ly * kLlfPitch + lx];
    const float thres = kZeroBias[c];
    const float* src = coef + c * nb * 64;
    int16_t* dst = ac + c * nb * 64;
    double err = 0, prev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int r8 = 0; r8 < 8; ++r8) {
#define PIKGPU_ENC_STEP(j)                                                   \
  {                                                                          \
    const int k = r8 * 8 + (j);                                              \
    const SubPos p = SubPosition(k, s, ix, iy, x0, y0, row_len);             \
    float v = k == 0 ? llf : src[p.at];                                      \
    if (c != 1) {                                                            \
      const float fac =                                                      \
          1.0f * float(cmap[(p.row >> 3) * P.tx + (p.col_block >> 3)] -      \
                       cmap_off) * cmap_scale;                               \
      v = v - fac * s_ydec[k][t];                                            \
    }                                                                        \
    const float val = v * (qm[p.pos] * qac);                                 \
    const int16_t q = DitherStep<j>(val, thres, k == 0, err, prev);          \
    dst[p.at] = q;                                                           \
    if (c == 1) s_ydec[k][t] = AdjustQuantBiasInt(1, q) * dm[p.pos] * inv_qac; \
  }
      PIKGPU_ENC_STEP(0) PIKGPU_ENC_STEP(1) PIKGPU_ENC_STEP(2)
      PIKGPU_ENC_STEP(3) PIKGPU_ENC_STEP(4) PIKGPU_ENC_STEP(5)
      PIKGPU_ENC_STEP(6) PIKGPU_ENC_STEP(7)
#undef PIKGPU_ENC_STEP
    }
  }
}

// ---------------------------------------------------------------------------
// Prediction residuals (ComputeCoefficients' "enc predictions",
// pik/compressed_image.cc:373-388), between the LLF / LF prediction the
// decoder's kernels form from dc_dec (the same ComputeLlf and PredictLf the
// encoder calls) and KEncQuant.

// PredictLfForEncoder's per-block step (ComputeDecoderBlockAnd2x2DC,
// pik/ac_predictions.cc:136-199, with ComputeDecoderCoefficients, :73-113)
// for one 8x8 dither sub-block (ix, iy) of a transform, one thread per
// storage block: channel order Y, X, B. With predict_lf the LF positions of
// `coef` become residuals against the LF prediction. The sub-block's LF
// positions k = 1, 8, 9 are quantized as QuantizeBlockAC will quantize them
// (their rounding errors depend only on k = 0, 1, 8, 9, so the rest of the
// block, still without its HF residual here, does not matter) and
// dequantised with Y's colour correlation; with the LF prediction added
// back (and the LLF position restored) they are the decoder-side corner
// values that DC2x2FromLowFrequencies reads. They go to `corner` (the
// coefficient layout), for KPred2x2Interior<true>.
__global__ void __launch_bounds__(256)
    KEncLf(FrameParams P, FrameBuffers B, float* coef, const float* itable,
           float gsf, float* corner) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.bx * P.by) return;
  const int y = i / P.bx, x = i - y * P.bx;
  const uint8_t a = B.acs[i];
  const int raw = a >> 4, idx = a & 15, s = AcsCovered(raw);
  const int iy = idx / s, ix = idx - iy * s;
  const int x0 = x - ix, y0 = y - iy;
  const int quant = B.qf[size_t(y0) * P.bx + x0];
  const int cf = B.dcf[(y0 >> 3) * P.tx + (x0 >> 3)];
  const int table = B.dqmap[cf * 256 + quant];
  const int kind = AcsQuantKind(raw);
  const float qac = gsf * quant;
  const float inv_qac = P.inv_global_scale / quant;
  const size_t nb = size_t(P.bx) * P.by;
  const size_t base = size_t(y0) * P.bx * 64 + size_t(x0) * 64;
  const int g = (y0 / kGroupDimBlocks) * P.gx + x0 / kGroupDimBlocks;
  const int lx0 = x0 % kGroupDimBlocks + 1, ly0 = y0 % kGroupDimBlocks + 1;
  const int ctile = (y0 >> 3) * P.tx + (x0 >> 3);
  constexpr float kZeroBias[3] = {0.65f, 0.6f, 0.7f};
  // F-matrix coordinates of the sub-block's LLF (k = 0) and LF positions
  // (k = 1, 8, 9): F[s*(k/8) + iy][s*(k%8) + ix]
  const int fi[4] = {iy, iy, s + iy, s + iy};
  const int fj[4] = {ix, s + ix, ix, s + ix};
  float ydec[4];
#pragma unroll 1
  for (int ci = 0; ci < 3; ++ci) {
    const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
    const uint32_t off = B.offsets[(table * kNumQuantKinds + kind) * 3 + c];
    const float* dm = B.table + off;
    const float* qm = itable + off;
    const float fac =
        c == 0 ? 1.0f * float(B.ytox[ctile] - 128) * (1.0f / 256)
               : 1.0f * float(B.ytob[ctile] - 0) * (1.0f / 128);
    const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane +
                      2 * ly0 * kG2Pitch + 2 * lx0;
    const float llf =
        B.llf[(size_t(g) * 3 + c) * kLlfPlane + (ly0 + iy) * kLlfPitch + lx0 + ix];
    float* cp = coef + c * nb * 64 + base;
    float dec[4];
    double err = 0, prev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int pos = fi[m] * 8 * s + fj[m];
      float v = m == 0 ? llf : cp[pos];
      if (m > 0 && P.predict_lf) {
        v -= lf[fi[m] * kG2Pitch + fj[m]];
        cp[pos] = v;  // the residual the block's quantization sees
      }
      if (c != 1) v -= ydec[m] * fac;
      const float val = v * (qm[pos] * qac);
      // k = 0, 1 are steps 0 and 1 of the first row, k = 8, 9 those of the
      // second (the steps between leave err and prev[0..1] as found)
      int16_t q;
      if (m == 0) {
        q = DitherStep<0>(val, kZeroBias[c], true, err, prev);
      } else if (m == 1) {
        q = DitherStep<1>(val, kZeroBias[c], false, err, prev);
      } else if (m == 2) {
        q = DitherStep<0>(val, kZeroBias[c], false, err, prev);
      } else {
        q = DitherStep<1>(val, kZeroBias[c], false, err, prev);
      }
      if (m == 0) {
        // the LLF position keeps its value (through the colour correlation
        // taken out and put back for X and B)
        dec[0] = c != 1 ? v + ydec[0] * fac : v;
      } else {
        const float d = AdjustQuantBiasInt(c, q) * dm[pos] * inv_qac;
        dec[m] = c != 1 ? d + ydec[m] * fac : d;
      }
    }
    if (c == 1) {
#pragma unroll
      for (int m = 0; m < 4; ++m) ydec[m] = dec[m];
    }
    float* out = corner + c * nb * 64 + base;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float v = dec[m];
      if (P.predict_lf) v = m == 0 ? llf : v + lf[fi[m] * kG2Pitch + fj[m]];
      out[fi[m] * 8 * s + fj[m]] = v;
    }
  }
}

// ComputePredictionResiduals (pik/ac_predictions.cc:637-650): per 32x32
// tile, the HF prediction the decoder adds (UpSample4x4BlurDCT, the blur of
// the tile's pred2x2 window and its forward transform, as in KBlocks)
// subtracted from the coefficients outside each transform's 2s x 2s
// corner.
__global__ void __launch_bounds__(256)
    KEncHfResidual(FrameParams P, FrameBuffers B, float* coef) {
  __shared__ float s_coef[3][kCoefPlane];
  __shared__ float s_pred[3][kTilePlane];
  __shared__ float s_tmp[3][kTilePlane];
  __shared__ uint8_t s_acs[16];
  __shared__ TileTasks s_fwd;
  const int tid = threadIdx.x;
  const int tiles_x = (P.bx + 3) / 4;
  const int tile = XcdTile(blockIdx.x, tiles_x * ((P.by + 3) / 4));
  if (tile < 0) return;
  const int tbx = (tile % tiles_x) * 4, tby = (tile / tiles_x) * 4;
  float (*s_p2)[100] = reinterpret_cast<float (*)[100]>(&s_tmp[0][0]);
  const int gxi = tbx / kGroupDimBlocks, gyi = tby / kGroupDimBlocks;
  const int g = gyi * P.gx + gxi;
  const int gbx0 = gxi * kGroupDimBlocks, gby0 = gyi * kGroupDimBlocks;
  const int p2r0 = 2 * (tby - gby0) + 1, p2c0 = 2 * (tbx - gbx0) + 1;
  const size_t nb = size_t(P.bx) * P.by, row = size_t(P.bx) * 64;
  if (tid < 16) {
    const int bxa = tbx + (tid & 3), bya = tby + (tid >> 2);
    s_acs[tid] = (bxa < P.bx && bya < P.by) ? B.acs[bya * P.bx + bxa]
                                            : uint8_t(0xFF);
  }
  for (int k = tid; k < 300; k += 256) {
    const int c = k / 100, i = (k % 100) / 10, j = k % 10;
    s_p2[c][k % 100] = B.pred2x2[(size_t(g) * 3 + c) * kG2Plane +
                                 (p2r0 + i) * kG2Pitch + p2c0 + j];
  }
  for (int k = tid; k < 3 * 16 * 64; k += 256) {
    const int c = k >> 10, blk = (k >> 6) & 15, e = k & 63;
    const int r = blk >> 2, b = blk & 3;
    const int yy = tby + r, xx = tbx + b;
    if (xx < P.bx && yy < P.by) {
      s_coef[c][CoefAt(r, b * 64 + e)] =
          coef[c * nb * 64 + size_t(yy) * row + size_t(xx) * 64 + e];
    }
  }
  __syncthreads();
  BuildTasks(s_acs, true, &s_fwd);
  const int pr = tid >> 3, pc0 = (tid & 7) << 2;
  const uint8_t pa = s_acs[(pr >> 3) * 4 + (pc0 >> 3)];
  if (pa != 0xFF && AcsPredictHF(pa >> 4)) {
    const int py = pr & 7, iy = py & 3;
    const int wr = 2 * (pr >> 3) + (py < 4 ? 0 : 1);
    const int wc = pc0 >> 2;
    for (int c = 0; c < 3; ++c) {
      float rw[3][3];
      for (int t = 0; t < 3; ++t) {
        for (int u = 0; u < 3; ++u) rw[t][u] = s_p2[c][(wr + t) * 10 + wc + u];
      }
      for (int j = 0; j < 4; ++j) {
        float v[3];
        for (int t = 0; t < 3; ++t) {
          v[t] = rw[t][0] * P.ub4[0][j] + rw[t][1] * P.ub4[1][j] +
                 rw[t][2] * P.ub4[2][j];
        }
        s_pred[c][PixAt(pr, pc0 + j)] =
            v[0] * P.ub4[0][iy] + v[1] * P.ub4[1][iy] + v[2] * P.ub4[2][iy];
      }
    }
  }
  __syncthreads();
  TileForward3<false, true, true>(s_fwd, &s_pred[0][0], &s_coef[0][0],
                                  &s_tmp[0][0]);
  __syncthreads();
  for (int k = tid; k < 3 * 16 * 64; k += 256) {
    const int c = k >> 10, blk = (k >> 6) & 15, e = k & 63;
    const int r = blk >> 2, b = blk & 3;
    const int yy = tby + r, xx = tbx + b;
    if (xx < P.bx && yy < P.by) {
      coef[c * nb * 64 + size_t(yy) * row + size_t(xx) * 64 + e] =
          s_coef[c][CoefAt(r, b * 64 + e)];
    }
  }
}

}  // namespace pikgpu
