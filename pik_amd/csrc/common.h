// Shared definitions of the gfx950 reconstruction kernels and their host
// set-up code.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#define __host__
#define __device__
#endif

namespace pikgpu {

constexpr int kBlockDim = 8;
constexpr int kGroupDimBlocks = 64;   // 512-px decode groups (pik/common.h:40)
constexpr int kTileDimBlocks = 8;     // 64-px tiles / colour tiles
constexpr int kEpfBorder = 6;         // pik/epf_target.cc:30
constexpr int kSigmaShift = 2;        // pik/epf.h:64
constexpr int kMinSigma = 4 << kSigmaShift;
constexpr int kMaxSigma = 168 << kSigmaShift;
constexpr int kClampedSAD = 507;      // pik/epf_target.cc:262
constexpr int kNumQuantKinds = 7;

// Raw AC strategy types (pik/ac_strategy.h:55-75).
enum Acs : int {
  kDCT = 0, kIdentity = 1, kDCT2X2 = 2, kDCT4X4 = 3, kDCT16 = 4, kDCT32 = 5,
  kLines = 6, kDCTNoHF = 7, kDCT4X4NoHF = 8
};

__host__ __device__ inline int AcsCovered(int raw) {
  return raw == kDCT32 ? 4 : (raw == kDCT16 ? 2 : 1);
}
// GetQuantKind (pik/ac_strategy.h:97-117).
__host__ __device__ inline int AcsQuantKind(int raw) {
  if (raw == kDCTNoHF) return 0;
  if (raw == kDCT4X4NoHF) return 3;
  return raw;
}
// PredictHF (pik/ac_strategy.h:132-136).
__host__ __device__ inline bool AcsPredictHF(int raw) {
  return raw != kDCT2X2 && raw != kIdentity && raw != kDCTNoHF &&
         raw != kDCT4X4NoHF && raw != kLines;
}
// Strategies whose 8x8 (or larger) block is a plain separable DCT.
__host__ __device__ inline bool AcsIsDCT(int raw) {
  return raw == kDCT || raw == kDCTNoHF || raw == kDCT16 || raw == kDCT32;
}
__host__ __device__ inline bool AcsIsDCT4(int raw) {
  return raw == kDCT4X4 || raw == kDCT4X4NoHF;
}

// Mirror of out-of-range coordinates with the edge pixel repeated
// (pik/image_ops.h:217-227).
__host__ __device__ inline int Mirror(int x, int size) {
  while (x < 0 || x >= size) {
    if (x < 0) {
      x = -x - 1;
    } else {
      x = 2 * size - 1 - x;
    }
  }
  return x;
}

// Per-frame constants derived on the host and passed to every kernel by
// value (one struct, < 4 KiB, lives in the kernel argument segment).
// Output formats of the final kernel.
enum { kOutLinear = 1, kOutSrgb8 = 2 };

struct FrameParams {
  int xsize, ysize;          // pixels
  int bx, by;                // blocks
  int px, py;                // padded pixels (8*bx, 8*by)
  int tx, ty;                // 64-px tiles
  int gx, gy;                // 512-px groups
  int predict_lf, predict_hf;
  int gaborish;              // strength != 0
  int ar;                    // adaptive reconstruction on
  int use_sharpened;
  int num_dots;
  int out_flags;             // kOutLinear | kOutSrgb8
  int downsample;            // 1, 2, 4 or 8
  int num_passes;            // AC passes accumulated (1 + pikgpu_frame_add_pass)
  int grayscale;             // rebuild the X / B DC from Y (GrayXyb)
  int dot_tiles;             // 32x32 tiles touched by dots
  int dot_tiles_x;           // tiles per row (ceil(px / 32))
  int pred_end[4];           // ends of the first-block classes in pred_list:
                             // DCT8-like, 2x2-Hadamard, DCT16, DCT32
  float inv_global_scale;    // Quantizer::InvGlobalScale()
  float inv_quant_dc;        // Quantizer::inv_quant_dc() (decoder form)
  float mul_dc[3];           // DC dequant multipliers
  float dc_half_step[3];     // AdaptiveDCReconstruction clamp
  float ytox_dc, ytob_dc;    // CfL for DC
  int dc_qf;                 // QuantDC()/2, quant field of the DC EPF
  float gab_w[3];            // Gaborish weights (center, edge, corner)
  float lfblur_w[3];         // LFPredictionBlur weights
  float ub4[3][4];           // UpSample4x4BlurDCT kernel
  float neg_bias[3];         // -kOpsinAbsorbanceBias (pik/opsin_params.h:36-43)
};

// Host-side computation of FrameParams' derived tables.
struct HostTables {
  int mul_table[kMaxSigma + 1];   // MulTable (pik/epf_target.cc:379-421)
};

void ComputeHostTables(HostTables* t);

// GrayXyb (pik/compressed_image.cc:94-209): 16 line pieces mapping a
// grayscale Y to X and B, fitted by linear regression over 1024 gray
// levels; index i = (int)((y - sub) * mul * 16), clamped to 0..15.
constexpr int kGrayPieces = 16;
struct GrayXybTable {
  double sub, mul;
  double x_slope[kGrayPieces], x_const[kGrayPieces];
  double b_slope[kGrayPieces], b_const[kGrayPieces];
};
void ComputeGrayXyb(GrayXybTable* t);
// LinearToXyb (pik/opsin_image.cc:34-47) with the reference's approximate
// cube root (pik/approx_cube_root.h).
void LinearToXybApprox(float r, float g, float b, float* x, float* y,
                       float* z);
// The library dequantisation tables (quant_tables.cpp); returns the number
// of floats, writes at most table_cap floats and offsets_cap offsets.
size_t DefaultDequantTables(float* table, size_t table_cap, uint32_t* offsets,
                            size_t offsets_cap, size_t* num_offsets);
void FillScaleTables(float* tscale, float* ascale);
// Squared norms D_k of the rows of the 1-D transposed-scaled IDCT basis for
// N = 4, 8, 16, 32, packed at AdjNormOffset(N) (the transpose of an inverse
// transform is its forward transform scaled by D_i D_j, see xform.h).
constexpr int kAdjNormFloats = 4 + 8 + 16 + 32;
__host__ __device__ inline int AdjNormOffset(int n) {
  return n == 4 ? 0 : (n == 8 ? 4 : (n == 16 ? 12 : 28));
}
void AdjointNorms(float* d);
void GaborishWeights(int strength, float w[3]);
void LfBlurWeights(float w[3]);
void Ub4Kernel(float k[3][4]);

}  // namespace pikgpu
