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

__device__ __forceinline__ void KBlocksBody(
    const FrameParams& P, const FrameBuffers& B, const DeviceConsts* K,
    const VBlock& vb) {
  // ~39 KB: four workgroups per CU. The IDCT's pixels reuse the HF
  // prediction's pixel buffer (dead once its transform is added).
  __shared__ float s_coef[3][kCoefPlane];
  __shared__ float s_pred[3][kTilePlane];
  __shared__ float s_tmp[3][kTilePlane];
  float* s_pix = &s_pred[0][0];
  __shared__ uint8_t s_acs[16];
  __shared__ TileTasks s_inv, s_fwd;

  const int tid = threadIdx.x;
  const int tiles_x = (P.bx + 3) / 4;
  const int tile = XcdTile(vb.x, tiles_x * ((P.by + 3) / 4));
  if (tile < 0) return;
  const int tbx = (tile % tiles_x) * 4, tby = (tile / tiles_x) * 4;  // blocks
  // pred2x2 window of the tile's HF prediction, 10 x 10 values per channel
  // (the tile's 8 x 8 plus the blur's one-value ring). It lives in the
  // transform scratch, which is first written after the last read of the
  // window (three workgroups per CU need the LDS under 160 KB / 3).
  float (*s_p2)[100] = reinterpret_cast<float (*)[100]>(&s_tmp[0][0]);
  const int gxi = tbx / kGroupDimBlocks, gyi = tby / kGroupDimBlocks;
  const int g = gyi * P.gx + gxi;
  const int gbx0 = gxi * kGroupDimBlocks, gby0 = gyi * kGroupDimBlocks;
  // pred2x2 rows / columns 2 * (tile's group-local block) + 1 .. + 10
  const int p2r0 = 2 * (tby - gby0) + 1, p2c0 = 2 * (tbx - gbx0) + 1;
  // Coefficient entries owned by this thread: storage row crow of the tile,
  // columns ccol..ccol+3 (one 8x8 storage block per 16 threads).
  const int crow = tid >> 6, ccol = (tid & 63) << 2;
  const int cbx = tbx + (ccol >> 6), cby = tby + crow;  // frame block
  const bool cvalid = cbx < P.bx && cby < P.by;
  const size_t plane = size_t(P.by) * P.bx * 64;
  const size_t at = size_t(cby) * P.bx * 64 + size_t(cbx) * 64 + (ccol & 63);
  const int ctile = (cby >> 3) * P.tx + (cbx >> 3);

  // Every global load that depends only on the tile's position is issued
  // before the first wait: the strategy bytes, the pred2x2 window, the
  // block's dequantisation descriptor (matrix slots and inv_global_scale /
  // qf, MakeBlockDq's values), its quantized coefficients and CfL factors.
  uint8_t a_own = 0xFF;
  if (tid < 16) {
    const int bxa = tbx + (tid & 3), bya = tby + (tid >> 2);
    if (bxa < P.bx && bya < P.by) a_own = B.acs[bya * P.bx + bxa];
  }
  float p2v[2] = {0.0f, 0.0f};
  if (P.predict_hf) {
    for (int u = 0; u < 2; ++u) {
      const int k = tid + 256 * u;
      if (k < 300) {
        const int c = k / 100, i = (k % 100) / 10, j = k % 10;
        p2v[u] = B.pred2x2[(size_t(g) * 3 + c) * kG2Plane +
                           (p2r0 + i) * kG2Pitch + p2c0 + j];
      }
    }
  }
  uint4 dsc = make_uint4(0, 0, 0, 0);
  uint2 q4[3] = {make_uint2(0, 0), make_uint2(0, 0), make_uint2(0, 0)};
  int32_t ytox = 128, ytob = 0;
  if (cvalid) {
    dsc = B.desc[2 * (cby * P.bx + cbx)];
    // four int16 coefficients, 8-byte aligned (at is a multiple of 4)
    for (int c = 0; c < 3; ++c) {
      q4[c] = *reinterpret_cast<const uint2*>(B.ac + c * plane + at);
    }
    ytox = B.ytox[ctile];
    ytob = B.ytob[ctile];
  }
  if (tid < 16) s_acs[tid] = a_own;
  if (P.predict_hf) {
    for (int u = 0; u < 2; ++u) {
      const int k = tid + 256 * u;
      if (k < 300) s_p2[k / 100][k % 100] = p2v[u];
    }
  }
  __syncthreads();
  // F-matrix coordinates of this thread's entries.
  const uint8_t ca = s_acs[crow * 4 + (ccol >> 6)];
  const int craw = ca >> 4, cidx = ca & 15;
  const int cs = ca == 0xFF ? 1 : AcsCovered(craw);
  const int lgs = cs >> 1;  // log2(cs): cs is 1, 2 or 4
  const int coby = crow - (cidx >> lgs), cobx = (ccol >> 6) - (cidx & (cs - 1));
  const int cflat0 = (crow - coby) * 64 * cs + (ccol - cobx * 64);
  const int cN = 8 * cs;
  // group-local origin block of the transform
  const int obxg = tbx + cobx - gbx0, obyg = tby + coby - gby0;
  // Entry j is an LLF position (its value comes from the LLF cache), an LF
  // position (the LF prediction is added), or neither.
  int fix[4];
  for (int j = 0; j < 4; ++j) {
    const int flat = cflat0 + j;
    const int fi = flat >> (3 + lgs), fj = flat & (cN - 1);
    fix[j] = (fi < cs && fj < cs)
                 ? 1
                 : ((P.predict_lf && fi < 2 * cs && fj < 2 * cs) ? 2 : 0);
  }

  // The dequant matrix entries need the descriptor, the LLF / LF values
  // the entries' positions: both go out before the task tables are built.
  float mv[3][4], fixv[3][4];
  if (cvalid) {
    const uint32_t slot[3] = {dsc.x, dsc.y, dsc.z};
    for (int c = 0; c < 3; ++c) {
      for (int j = 0; j < 4; ++j) mv[c][j] = B.table[slot[c] + (ccol & 63) + j];
    }
    for (int c = 0; c < 3; ++c) {
      const float* llf = B.llf + (size_t(g) * 3 + c) * kLlfPlane;
      const float* lf = B.lf2x2 + (size_t(g) * 3 + c) * kG2Plane;
      for (int j = 0; j < 4; ++j) {
        const int flat = cflat0 + j;
        const int fi = flat >> (3 + lgs), fj = flat & (cN - 1);
        fixv[c][j] = 0.0f;
        if (fix[j] == 1) {
          fixv[c][j] = llf[(obyg + 1 + fi) * kLlfPitch + obxg + 1 + fj];
        } else if (fix[j] == 2) {
          fixv[c][j] = lf[(2 * (obyg + 1) + fi) * kG2Pitch + 2 * (obxg + 1) + fj];
        }
      }
    }
  }
  BuildTasks(s_acs, false, &s_inv);
  if (P.predict_hf) BuildTasks(s_acs, true, &s_fwd);

  // 1. dequantise Y, then X and B with their colour correlation
  //    (Dequant::DoAC, pik/compressed_image.cc:586-693), and 2. the LLF and
  //    LF predictions (UpdateLfForDecoder, pik/ac_predictions.cc:560-620).
  if (cvalid) {
    const float scaled = __uint_as_float(dsc.w);
    const float ccf[3] = {float(ytox - 128) * 0.00390625f, 0.0f,
                          float(ytob) * 0.0078125f};
    // Every pass in order: the first stores, the later ones add their own
    // dequantised values (Dequant<false>::DoAC, pik/compressed_image.cc:
    // 695-703); then the LLF / LF positions.
    float acc[3][4];
    for (int p = 0; p < P.num_passes; ++p) {
      uint2 qp[3] = {q4[0], q4[1], q4[2]};
      if (p > 0) {
        for (int c = 0; c < 3; ++c) {
          qp[c] = *reinterpret_cast<const uint2*>(AcPlane(P, B, p, c) + at);
        }
      }
      int16_t qv[3][4];
      for (int c = 0; c < 3; ++c) {
        qv[c][0] = int16_t(qp[c].x & 0xFFFF);
        qv[c][1] = int16_t(qp[c].x >> 16);
        qv[c][2] = int16_t(qp[c].y & 0xFFFF);
        qv[c][3] = int16_t(qp[c].y >> 16);
      }
      float yv[4];
      for (int ci = 0; ci < 3; ++ci) {
        const int c = ci == 0 ? 1 : (ci == 1 ? 0 : 2);
        for (int j = 0; j < 4; ++j) {
          // DequantWith's operation order
          const float mul = mv[c][j] * scaled;
          const float d = AdjustQuantBias(c, float(qv[c][j])) * mul;
          const float v = c == 1 ? d : ccf[c] * yv[j] + d;
          if (c == 1) yv[j] = v;
          acc[c][j] = p == 0 ? v : v + acc[c][j];
        }
      }
    }
    for (int c = 0; c < 3; ++c) {
      for (int j = 0; j < 4; ++j) {
        float v = acc[c][j];
        if (fix[j] == 1) {
          v = fixv[c][j];
        } else if (fix[j] == 2) {
          v += fixv[c][j];
        }
        s_coef[c][CoefAt(crow, ccol + j)] = v;
      }
    }
  }
  // Pixel quad of this thread: row pr, columns pc0..pc0+3.
  const int pr = tid >> 3, pc0 = (tid & 7) << 2;
  // 3. HF prediction (AddPredictions / UpSample4x4BlurDCT<true>,
  //    pik/ac_predictions.cc:263-374): 4x nearest-neighbour upsampling of
  //    pred2x2 with the separable 3-tap blur, forward transform, added to
  //    all but the 2s x 2s low-frequency corner.
  if (P.predict_hf) {
    const uint8_t pa = s_acs[(pr >> 3) * 4 + (pc0 >> 3)];
    if (pa != 0xFF && AcsPredictHF(pa >> 4)) {
      // Pixel row pr of the tile reads pred2x2 rows rbase + 1 .. + 3 and
      // its four columns (one upsampled pred2x2 column, q = j) columns
      // xq + 1 .. + 3, where rbase = 2 * block row + (pr & 7 >= 4) and
      // xq = pc0 / 4, both tile-local; in window coordinates that is rows
      // rbase - 1 + 1 + t and columns xq + t.
      const int py = pr & 7, iy = py & 3;
      const int wr = 2 * (pr >> 3) + (py < 4 ? 0 : 1);  // window row of t = 0
      const int wc = pc0 >> 2;                         // window column
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
    // forward transform of the prediction, added to the coefficients
    // outside each transform's low-frequency corner
    TileForward3<false, true>(s_fwd, &s_pred[0][0], &s_coef[0][0],
                              &s_tmp[0][0]);
  }
  __syncthreads();
  if (B.keep && cvalid) {
    for (int c = 0; c < 3; ++c) {
      float* dst = B.coeffs + (size_t(c) * P.by + cby) * (P.bx * 64) + cbx * 64;
      for (int j = 0; j < 4; ++j) dst[(ccol + j) & 63] = s_coef[c][CoefAt(crow, ccol + j)];
    }
  }
  // 4. IDCT (InverseIntegralTransform, pik/compressed_image.cc:869-895)
  TileInverse3(s_inv, &s_coef[0][0], s_pix, &s_tmp[0][0]);
  if (P.downsample == 1) {
    const int py = tby * 8 + pr;
    if (py < P.py) {
      for (int c = 0; c < 3; ++c) {
        float* dst = B.idct + (size_t(c) * P.py + py) * B.ipitch + tbx * 8 + pc0;
        for (int j = 0; j < 4; ++j) {
          if (tbx * 8 + pc0 + j < P.px) {
            dst[j] = s_pix[c * kTilePlane + PixAt(pr, pc0 + j)];
          }
        }
      }
    }
    return;
  }
  // Downsampled decode (InverseIntegralTransform with downsample,
  // pik/compressed_image.cc:896-933): each ds x ds square of a transform's
  // pixels (transforms are 8-aligned, so the squares are tile-aligned) is
  // summed in row order and scaled by 1 / ds^2, into planes of
  // (py / ds) x ipitch.
  const int ds = P.downsample, n = 32 / ds;
  const float mean_mul = 1.0f / float(ds * ds);
  const int ox0 = tbx * 8 / ds, oy0 = tby * 8 / ds;
  const int opx = P.px / ds, opy = P.py / ds;
  const size_t oplane = size_t(opy) * B.ipitch;
  for (int k = tid; k < 3 * n * n; k += 256) {
    const int c = k / (n * n), rem = k - c * (n * n);
    const int oy = rem / n, ox = rem - oy * n;
    if (ox0 + ox >= opx || oy0 + oy >= opy) continue;
    float sum = 0.0f;
    for (int iy = 0; iy < ds; ++iy) {
      for (int ix = 0; ix < ds; ++ix) {
        sum += s_pix[c * kTilePlane + PixAt(oy * ds + iy, ox * ds + ix)];
      }
    }
    B.idct[c * oplane + size_t(oy0 + oy) * B.ipitch + ox0 + ox] =
        sum * mean_mul;
  }
}
__global__ void __launch_bounds__(256, 4)
    KBlocks(FrameParams P, FrameBuffers B, const DeviceConsts* K) {
  KBlocksBody(P, B, K, PlainBlock());
}

// One launch over every frame of a batch (BatchSpan).
__global__ void __launch_bounds__(256, 4) KBlocksBatch(PIKGPU_BATCH_PARAMS) {
  PIKGPU_BATCH_ONE(KBlocksBody(P, B, K, vb));
}

}  // namespace pikgpu
