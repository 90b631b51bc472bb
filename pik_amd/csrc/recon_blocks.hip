// Block kernel: dequantisation + CfL, lowest/low-frequency predictions, HF
// prediction and the variable-size IDCT, one 32x32-pixel tile per
// workgroup (gfx950, wave64, 256 threads = 4 waves).
//
// Follows, per tile and plane (Y first so X/B can apply CfL):
//   Dequant<true>::DoAC                pik/compressed_image.cc:586-693
//   UpdateLfForDecoder (LLF, LF add)   pik/ac_predictions.cc:560-620
//   AddPredictions/UpSample4x4BlurDCT  pik/ac_predictions.cc:263-374, 652-665
//   InverseIntegralTransform           pik/compressed_image.cc:869-895
//   AcStrategy::TransformToPixels      pik/ac_strategy.cc:674-840
// Tiles are 32-pixel aligned, so every transform (at most 32x32, aligned to
// its size) lies inside one tile and every tile inside one 512-px group.
#include "kernels.h"

namespace pikgpu {

__global__ void __launch_bounds__(256)
    KBlocks(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  __shared__ float s_coef[1024];
  __shared__ float s_y[1024];
  __shared__ float s_pix[1024];
  __shared__ float s_pred[1024];
  __shared__ float s_tmp[1024];
  __shared__ uint8_t s_acs[16];

  const int tid = threadIdx.x;
  const int tbx = blockIdx.x * 4, tby = blockIdx.y * 4;  // tile origin, blocks
  if (tid < 16) {
    const int bxa = tbx + (tid & 3), bya = tby + (tid >> 2);
    s_acs[tid] = (bxa < P.bx && bya < P.by) ? B.acs[bya * P.bx + bxa] : 0xFF;
  }
  __syncthreads();
  const TileXform X = MakeXform(s_acs, tid);

  const int gxi = tbx / kGroupDimBlocks, gyi = tby / kGroupDimBlocks;
  const int g = gyi * P.gx + gxi;
  const int gbx0 = gxi * kGroupDimBlocks, gby0 = gyi * kGroupDimBlocks;

  // Coefficient entries owned by this thread: row crow, cols ccol..ccol+3.
  const int crow = tid >> 6, ccol = (tid & 63) << 2;
  const int cbx = tbx + (ccol >> 6), cby = tby + crow;  // frame block
  const bool cvalid = cbx < P.bx && cby < P.by;
  // F-matrix coordinates of those entries.
  const uint8_t ca = s_acs[crow * 4 + (ccol >> 6)];
  const int craw = ca >> 4, cidx = ca & 15;
  const int cs = ca == 0xFF ? 1 : AcsCovered(craw);
  const int coby = crow - cidx / cs, cobx = (ccol >> 6) - cidx % cs;
  const int cflat0 = (crow - coby) * 64 * cs + (ccol - cobx * 64);
  const int cN = 8 * cs;
  // group-local origin block of the transform
  const int obxg = tbx + cobx - gbx0, obyg = tby + coby - gby0;

  for (int ci = 0; ci < 3; ++ci) {
    const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
    // 1. dequantise (+ CfL)
    if (cvalid) {
      for (int j = 0; j < 4; ++j) {
        const int col = cbx * 64 + ((ccol + j) & 63);
        const float yv = c == 1 ? 0.0f : s_y[crow * 256 + ccol + j];
        const float v = DequantCoef(P, B, c, cby, col, yv);
        s_coef[crow * 256 + ccol + j] = v;
        if (c == 1) s_y[crow * 256 + ccol + j] = v;
      }
    }
    // 2. LLF and LF predictions
    if (cvalid) {
      const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
      const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
      for (int j = 0; j < 4; ++j) {
        const int flat = cflat0 + j;
        const int fi = flat / cN, fj = flat % cN;
        float& v = s_coef[crow * 256 + ccol + j];
        if (fi < cs && fj < cs) {
          v = llf[(obyg + 1 + fi) * kLlfPitch + obxg + 1 + fj];
        } else if (P.predict_lf && fi < 2 * cs && fj < 2 * cs) {
          v += lf[(2 * (obyg + 1) + fi) * kG2Pitch + 2 * (obxg + 1) + fj];
        }
      }
    }
    __syncthreads();
    // 3. HF prediction: blurred 4x upsampling of pred2x2, forward
    //    transform, added except to the LF/LLF corner.
    if (P.predict_hf) {
      // strategy of the block holding this thread's pixel quad
      const uint8_t pa = s_acs[(X.r >> 3) * 4 + (X.c0 >> 3)];
      if (pa != 0xFF && AcsPredictHF(pa >> 4)) {
        const float* p2 = B.pred2x2 + (size_t(g) * 3 + c) * kG2Plane;
        const int by_l = (tby - gby0) + (X.r >> 3);  // group-local block row
        const int R0 = 2 * by_l;
        const int py = X.r & 7;
        const int iy = py & 3;
        const int rbase = py < 4 ? R0 : R0 + 1;
        for (int j = 0; j < 4; ++j) {
          const int Xc = (tbx - gbx0) * 8 + X.c0 + j;  // group-local column
          const int xq = Xc >> 2, q = Xc & 3;
          float v[3];
          for (int t = 0; t < 3; ++t) {
            const float* row = p2 + (rbase + t + 1) * kG2Pitch;
            v[t] = row[xq + 1] * P.ub4[0][q] + row[xq + 2] * P.ub4[1][q] +
                   row[xq + 3] * P.ub4[2][q];
          }
          s_pred[X.r * 32 + X.c0 + j] =
              v[0] * P.ub4[0][iy] + v[1] * P.ub4[1][iy] + v[2] * P.ub4[2][iy];
        }
      }
      __syncthreads();
      TileForward(K, X, s_pred, s_pix, s_tmp, true);
      // s_pix now holds the prediction's coefficients for PredictHF blocks.
      if (cvalid && ca != 0xFF && AcsPredictHF(craw)) {
        for (int j = 0; j < 4; ++j) {
          const int flat = cflat0 + j;
          const int fi = flat / cN, fj = flat % cN;
          if (!(fi < 2 * cs && fj < 2 * cs)) {
            s_coef[crow * 256 + ccol + j] += s_pix[crow * 256 + ccol + j];
          }
        }
      }
      __syncthreads();
    }
    if (B.keep && cvalid) {
      float* dst = B.coeffs + (size_t(c) * P.by + cby) * (P.bx * 64) + cbx * 64;
      for (int j = 0; j < 4; ++j) dst[(ccol + j) & 63] = s_coef[crow * 256 + ccol + j];
    }
    // 4. IDCT
    TileInverse(K, X, s_coef, s_pix, s_tmp);
    const int py = tby * 8 + X.r;
    if (py < P.py) {
      float* dst = B.idct + (size_t(c) * P.py + py) * B.ipitch + tbx * 8 + X.c0;
      for (int j = 0; j < 4; ++j) {
        if (tbx * 8 + X.c0 + j < P.px) dst[j] = s_pix[X.r * 32 + X.c0 + j];
      }
    }
    __syncthreads();
  }
}

}  // namespace pikgpu
