// Device-side data layout shared by the reconstruction kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"
#include "../../include/pikgpu.h"

namespace pikgpu {

// Per-context constant tables in device memory (set up once).
struct DeviceConsts {
  float tscale[5][8][8];         // DCTTotalScale<2^(i+1)>(x, y) at [i][y][x]
  float ascale[5][8][8];         // DCTInvTotalScale<2^(i+1)>(x, y)
  int mul_table[kMaxSigma + 1];  // EPF WeightFast multipliers
  float opsin_inv[9];            // inverse opsin absorbance matrix
  float adj_norm[kAdjNormFloats];  // AdjointNorms
  GrayXybTable gray;               // ComputeGrayXyb
};

// LF-domain per-group caches: (gw+2) x (gh+2) values per plane for llf and
// 2x that in each direction for the 2x grids, at fixed pitches.
constexpr int kLlfPitch = kGroupDimBlocks + 2;       // 66
constexpr int kG2Pitch = 2 * (kGroupDimBlocks + 2);  // 132
constexpr int kLlfPlane = kLlfPitch * kLlfPitch;
constexpr int kG2Plane = kG2Pitch * kG2Pitch;

// Device pointers and sizes of one frame.
struct FrameBuffers {
  // inputs
  const int16_t* ac;      // 3 planes, each by x (bx*64)
  const int16_t* ac_extra;  // passes 1 .. num_passes - 1, 3 planes each
  const int16_t* dc16;    // 3 planes, each by x bx
  const int32_t* qf;      // by x bx
  const uint8_t* acs;     // by x bx
  const uint8_t* lut;     // by x bx
  const uint8_t* dcf;     // ty x tx
  const int32_t* ytox;    // ty x tx
  const int32_t* ytob;    // ty x tx
  const uint8_t* dqmap;   // 16 x 256
  const float* table;
  const uint32_t* offsets;
  int num_tables;
  const pikgpu_dot* dots;
  // dots binned by 32x32 tile: tile ids, CSR starts (dot_tiles + 1) and the
  // dot indices of each tile in application order
  const int32_t* dot_tile_ids;
  const int32_t* dot_tile_start;
  const int32_t* dot_list;
  // first blocks grouped by the class of their pred2x2 computation (block
  // indices; classes end at FrameParams::pred_end)
  const uint32_t* pred_list;
  // intermediates
  // Per block two descriptors (KDcDequant): {table offset of the block's
  // dequant matrix slot for channels 0..2 (offsets[(qt * 7 + kind) * 3 + c]
  // + idx * 64), bits of inv_global_scale / qf} for the block's own quant
  // value, then the same for the quant value of its transform's first
  // block (AR's quantisation intervals).
  uint4* desc;
  float* dcq;             // 3 planes by x bx (DequantDC)
  float* dc;              // 3 planes by x bx (after DC AR)
  float* llf;             // [group][3][kLlfPlane]
  float* tmp2x2;          // [group][3][kG2Plane]
  float* lf2x2;           // [group][3][kG2Plane]
  float* pred2x2;         // [group][3][kG2Plane]
  float* coeffs;          // 3 planes by x (bx*64) (stage dump only)
  float* idct;            // 3 planes py x ipitch
  float* guide;           // stage dump only, same layout as idct
  float* epf;             // stage dump only
  float* ar;              // 3 planes py x ipitch
  // u8 EPF guide, 3 planes of grows rows x gpitch bytes; pixel (x, y) at
  // (x + 6, y + 6), the 6-pixel border mirrored (KGuide, KGuideRing)
  uint8_t* gimg;
  int gpitch, grows;
  float* dots_img;        // stage dump only
  float* gab;             // stage dump only
  float* out;             // 3 planes ysize x opitch
  uint8_t* srgb8;         // ysize rows of spitch bytes, interleaved RGB
  float* mm_partial;      // [num_wg][6]
  float* mm;              // [8]: min[3], max[3], stretch/mul, all_min
  int ipitch;             // floats per row of idct/ar
  int opitch;             // floats per row of out
  int spitch;             // bytes per row of srgb8
  int keep;               // write stage dumps
};

// Inputs and outputs of one forward + adjoint evaluation (device memory).
// Image planes: ysize rows of the given pitch (floats); block fields:
// ysize_blocks rows; coefficient planes: ysize_blocks rows of
// xsize_blocks * 64 values.
struct AdjArgs {
  const float* qf;           // real-valued quant field, > 0
  int qf_pitch;
  const float* target[3];    // linear target, or null (then upstream)
  const float* upstream[3];  // dL/dlinear when target is null
  int in_pitch;
  float* linear[3];          // optional forward output
  int lin_pitch;
  float* dqf;                // dL/dqf
  int dqf_pitch;
  float* dcoef[3];           // optional dL/dd (before CfL)
  size_t dc_pitch;
  float* loss_part;          // per-tile sums of squared error, or null
};

}  // namespace pikgpu
