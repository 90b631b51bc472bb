// Test harness around the google/pik reference, compiled together with the
// reference's own sources (read in place under the reference checkout) by
// oracle/Makefile into oracle/_ref/. TEST INFRASTRUCTURE ONLY: it is a
// checker for the MI355X path and the CPU baseline, never part of the
// product.
//
// It exposes the reference's decode-side reconstruction: pikref_recon runs
// DequantDC + AdaptiveDCReconstruction (pik/compressed_dc.cc:181-230,
// pik/adaptive_reconstruction.cc:300-332), then per 512x512 group
// InitializeDecCache / DequantImageAC / ReconOpsinImage as PikGroupToPixels
// and RoundtripImage do (pik/pik_frame.cc:846-928,
// pik/adaptive_quantization.cc:613-667), then the steps of
// FinalizeFrameDecoding (pik/compressed_image.cc:1110-1181), on a
// pikgpu_frame, with optional dumps of every intermediate stage; plus single
// transforms, Gaborish, OpsinToLinear and the EPF weight for unit checks.
//
// The reference encoder is not built: its AC-strategy search, dot detection
// and butteraugli quantizer search all link code that needs lcms2, which the
// checkout does not contain. Frames are made by the repository's own
// generator instead.

// The standard headers the reference uses are included first: the harness
// then opens the reference's private members (placements of the dot
// dictionary, the dequant offset table) to export them.
#include <emmintrin.h>
#include <inttypes.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>
#include <time.h>
#include <x86intrin.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#define private public
#include "pik/block_dictionary.h"
#include "pik/quant_weights.h"
#undef private

#include "pik/ac_strategy.h"
#include "pik/adaptive_quantization.h"
#include "pik/adaptive_reconstruction.h"
#include "pik/ar_control_field.h"
#include "pik/bilinear_transform.h"
#include "pik/bit_reader.h"
#include "pik/butteraugli_comparator.h"
#include "pik/color_correlation.h"
#include "pik/compressed_dc.h"
#include "pik/compressed_image.h"
#include "pik/data_parallel.h"
#include "pik/entropy_coder.h"
#include "pik/fields.h"
#include "pik/epf.h"
#include "pik/gaborish.h"
#include "pik/headers.h"
#include "pik/image.h"
#include "pik/noise.h"
#include "pik/opsin_image.h"
#include "pik/opsin_inverse.h"
#include "pik/padded_bytes.h"
#include "pik/pik_params.h"
#include "pik/quantizer.h"
#include "pik/simd/targets.h"
#include "pik/write_bits.h"

#include "../include/pikgpu.h"

namespace pik {
// Static dot dictionary, pik/block_dictionary.cc:26-29.
extern QuantizedBlock kStaticBlocks[];
}  // namespace pik

using namespace pik;

#define API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_error;

// FrameHeader::kGrayscaleOpt from either of the frame's two spellings.
bool IsGray(const pikgpu_frame* f) {
  return f->grayscale != 0 || (f->flags & PIKGPU_FLAG_GRAYSCALE_OPT) != 0;
}

int Fail(const char* msg) {
  g_error = msg;
  return -1;
}

std::unique_ptr<ThreadPool> MakePool(int num_threads) {
  if (num_threads <= 0) return nullptr;
  return std::unique_ptr<ThreadPool>(new ThreadPool(num_threads));
}

void InitEpfOnce() {
  static std::once_flag once;
  // Same initialisation as the decoder's CodecContext (pik/codec_impl.cc:188).
  std::call_once(once,
                 [] { TargetBitfield().Foreach(InitEdgePreservingFilter()); });
}

void ExportTables(const DequantMatrices& m, std::vector<float>* table,
                  std::vector<uint32_t>* offsets) {
  table->assign(m.table_, m.table_ + m.Size());
  offsets->assign(m.table_offsets_.begin(), m.table_offsets_.end());
}

BlockDictionary ImportDots(const pikgpu_frame& f) {
  std::vector<QuantizedBlock> dict;
  std::vector<BlockPosition> pos;
  for (size_t i = 0; i < f.num_dots; ++i) {
    const pikgpu_dot& d = f.dots[i];
    QuantizedBlock b;
    memset(&b, 0, sizeof(b));
    b.xsize = d.xsize;
    b.ysize = d.ysize;
    for (int c = 0; c < 3; ++c) memcpy(b.pixels[c], d.pixels[c], 64);
    dict.push_back(b);
    if (d.transform) {
      BlockPosition p(d.x, d.y, i, d.dx, d.dy, d.width);
      p.width = d.width;
      pos.push_back(p);
    } else {
      pos.push_back(BlockPosition(d.x, d.y, i));
    }
  }
  return BlockDictionary(dict, pos);
}

void DumpImage3F(const Image3F& img, float* const* planes, size_t xs,
                 size_t ys) {
  if (planes == nullptr) return;
  for (int c = 0; c < 3; ++c) {
    if (planes[c] == nullptr) continue;
    for (size_t y = 0; y < ys; ++y) {
      memcpy(planes[c] + y * xs, img.ConstPlaneRow(c, y), xs * sizeof(float));
    }
  }
}

// Quantizer as the decoder reconstructs it from its three serialized bytes
// (Quantizer::Decode, pik/quantizer.cc:63-71): inv_quant_dc is derived there.
Quantizer DecoderQuantizer(const DequantMatrices* m, size_t bx, size_t by,
                           int quant_dc, int global_scale) {
  Quantizer q(m, bx, by, quant_dc, global_scale);
  const std::string bytes = q.Encode(nullptr);
  PaddedBytes padded(bytes.size() + 8);
  memset(padded.data(), 0, padded.size());
  memcpy(padded.data(), bytes.data(), bytes.size());
  BitReader br(padded.data(), padded.size());
  q.Decode(&br);
  return q;
}

}  // namespace

API const char* pikref_last_error() { return g_error.c_str(); }

API int pikref_abi_version() { return PIKGPU_ABI_VERSION; }

// Default dequant tables (DequantMatrices with library encodings,
// pik/quant_weights.cc:607-926). Returns the number of floats/offsets; copies
// up to the given capacities.
API int pikref_default_tables(float* table, size_t table_cap,
                              uint32_t* offsets, size_t offsets_cap,
                              size_t* table_size, size_t* num_offsets) {
  DequantMatrices m(/*need_inv_matrices=*/false);
  std::vector<float> t;
  std::vector<uint32_t> o;
  ExportTables(m, &t, &o);
  if (table_size) *table_size = t.size();
  if (num_offsets) *num_offsets = o.size();
  if (table) memcpy(table, t.data(), std::min(t.size(), table_cap) * 4);
  if (offsets) memcpy(offsets, o.data(), std::min(o.size(), offsets_cap) * 4);
  return 0;
}

// DecodeFrame's step after FinalizeFrameDecoding for kGrayscaleOpt frames
// (pik/pik_frame.cc:1155-1166): DecodeFrame itself needs a bitstream, so the
// loop is restated here on the finalized image.
static void ForceGray(Image3F* img) {
  for (size_t y = 0; y < img->ysize(); ++y) {
    float* PIK_RESTRICT r = img->PlaneRow(0, y);
    float* PIK_RESTRICT g = img->PlaneRow(1, y);
    float* PIK_RESTRICT b = img->PlaneRow(2, y);
    for (size_t x = 0; x < img->xsize(); ++x) {
      const float gray = r[x] * 0.299 + g[x] * 0.587 + b[x] * 0.114;
      r[x] = g[x] = b[x] = gray;
    }
  }
}

// Reference reconstruction. `stages` may be NULL, or point to
// PIKGPU_NUM_STAGES entries of 3 plane pointers (each may be NULL) receiving
// dense images of the sizes documented in pikgpu.h. `out` receives the final
// linear image (xsize x ysize per plane, dense). Returns 0 on success.
// Downsample 2, 4 or 8 takes
// ReconOpsinImage's downsampled IDCT and FinalizeFrameDecoding's
// downsampled branch (no AR, no Gaborish, ShrinkTo the downsampled size).
static int Recon(const pikgpu_frame* f, int num_threads, size_t downsample,
                 float* const* stages, float* const* out, int num_extra = 0,
                 const int16_t* const* extra_ac = nullptr,
                 size_t extra_stride = 0) {
  if (f == nullptr) return Fail("null frame");
  if (downsample != 1 && downsample != 2 && downsample != 4 &&
      downsample != 8) {
    return Fail("downsample must be 1, 2, 4 or 8");
  }
  if (downsample != 1 && (stages != nullptr || f->num_dots != 0)) {
    return Fail("downsampled decode: no stage dumps, no dots");
  }
  InitEpfOnce();
  std::unique_ptr<ThreadPool> pool = MakePool(num_threads);
  const size_t xsize = f->xsize, ysize = f->ysize;
  const size_t bx = DivCeil<size_t>(xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(ysize, kBlockDim);
  auto stage = [&](int s) -> float* const* {
    return stages == nullptr ? nullptr : stages + 3 * s;
  };

  // Only the library tables can be instantiated here; check the frame uses
  // them.
  DequantMatrices matrices(/*need_inv_matrices=*/false);
  {
    std::vector<float> t;
    std::vector<uint32_t> o;
    ExportTables(matrices, &t, &o);
    if (f->dequant_table_size != t.size() ||
        memcmp(f->dequant_table, t.data(), t.size() * 4) != 0 ||
        f->num_dequant_offsets != o.size() ||
        memcmp(f->dequant_offsets, o.data(), o.size() * 4) != 0) {
      return Fail("frame does not use the default dequant tables");
    }
  }

  FrameHeader frame_header;
  frame_header.flags = IsGray(f) ? FrameHeader::kGrayscaleOpt : 0;
  frame_header.gaborish = GaborishStrength(f->gaborish);
  frame_header.predict_lf = f->predict_lf != 0;
  frame_header.predict_hf = f->predict_hf != 0;
  frame_header.have_adaptive_reconstruction = f->adaptive_reconstruction != 0;
  frame_header.epf_params.enable_adaptive = f->epf_enable_adaptive != 0;
  frame_header.epf_params.use_sharpened = f->epf_use_sharpened != 0;
  GroupHeader group_header;

  Quantizer quantizer =
      DecoderQuantizer(&matrices, bx, by, f->quant_dc, f->global_scale);
  {
    ImageI qf(bx, by);
    for (size_t y = 0; y < by; ++y) {
      memcpy(qf.Row(y), f->quant_field + y * f->quant_field_stride, bx * 4);
    }
    quantizer.SetRawQuantField(std::move(qf));
  }

  ColorCorrelationMap cmap(xsize, ysize);
  for (size_t y = 0; y < cmap.ytox_map.ysize(); ++y) {
    memcpy(cmap.ytox_map.Row(y), f->ytox_map + y * f->ytox_stride,
           cmap.ytox_map.xsize() * 4);
    memcpy(cmap.ytob_map.Row(y), f->ytob_map + y * f->ytob_stride,
           cmap.ytob_map.xsize() * 4);
  }
  cmap.ytox_dc = f->ytox_dc;
  cmap.ytob_dc = f->ytob_dc;

  FrameDecCache fdc;
  fdc.grayscale = IsGray(f);
  // AC strategy: the frame holds combined bytes; SetFromRaw expands raw
  // types (pik/ac_strategy.cc:1066-1099).
  {
    ImageB raw(bx, by);
    for (size_t y = 0; y < by; ++y) {
      for (size_t x = 0; x < bx; ++x) {
        raw.Row(y)[x] = f->ac_strategy[y * f->ac_strategy_stride + x] >> 4;
      }
    }
    fdc.ac_strategy = AcStrategyImage(bx, by);
    fdc.ac_strategy.SetFromRaw(Rect(0, 0, bx, by), raw);
    for (size_t y = 0; y < by; ++y) {
      if (memcmp(fdc.ac_strategy.ConstRaw().ConstRow(y),
                 f->ac_strategy + y * f->ac_strategy_stride, bx) != 0) {
        return Fail("inconsistent ac_strategy bytes");
      }
    }
  }
  fdc.raw_quant_field = CopyImage(quantizer.RawQuantField());
  fdc.ar_sigma_lut_ids = ImageB(bx, by);
  for (size_t y = 0; y < by; ++y) {
    memcpy(fdc.ar_sigma_lut_ids.Row(y), f->sigma_lut_ids + y * f->sigma_lut_stride,
           bx);
  }
  const size_t tx = DivCeil<size_t>(xsize, kColorTileDim);
  const size_t ty = DivCeil<size_t>(ysize, kColorTileDim);
  fdc.dequant_control_field = ImageB(tx, ty);
  for (size_t y = 0; y < ty; ++y) {
    memcpy(fdc.dequant_control_field.Row(y),
           f->dequant_cf + y * f->dequant_cf_stride, tx);
  }
  memcpy(fdc.dequant_map, f->dequant_map, sizeof(fdc.dequant_map));

  // DC: DequantDC + AdaptiveDCReconstruction (DecodeDCGroups,
  // pik/compressed_dc.cc:436-528).
  {
    float mul_dc[3];
    for (int c = 0; c < 3; ++c) {
      mul_dc[c] = quantizer.DequantMatrix(0, kQuantKindDCT8, c)[0] *
                  quantizer.inv_quant_dc();
    }
    const float ytox_dc = ColorCorrelationMap::YtoX(1.0f, cmap.ytox_dc);
    const float ytob_dc = ColorCorrelationMap::YtoB(1.0f, cmap.ytob_dc);
    Image3S dc16(bx, by);
    for (int c = 0; c < 3; ++c) {
      for (size_t y = 0; y < by; ++y) {
        memcpy(dc16.PlaneRow(c, y), f->dc[c] + y * f->dc_stride, bx * 2);
      }
    }
    // DequantDC (pik/compressed_dc.cc:181-230) sits in an anonymous
    // namespace of its TU; the same per-element arithmetic is done here:
    // Y = q * mul, X/B = ytox/ytob * Y + q * mul (mul_add of the portable
    // target is a separate multiply and add).
    fdc.dc = Image3F(bx, by);
    for (size_t y = 0; y < by; ++y) {
      const int16_t* qy = dc16.ConstPlaneRow(1, y);
      float* ry = fdc.dc.PlaneRow(1, y);
      for (size_t x = 0; x < bx; ++x) ry[x] = float(qy[x]) * mul_dc[1];
    }
    for (int c = 0; c < 3; c += 2) {
      const float ymul = c == 0 ? ytox_dc : ytob_dc;
      for (size_t y = 0; y < by; ++y) {
        const int16_t* q = dc16.ConstPlaneRow(c, y);
        const float* ry = fdc.dc.ConstPlaneRow(1, y);
        float* r = fdc.dc.PlaneRow(c, y);
        for (size_t x = 0; x < bx; ++x) {
          const float dq = float(q[x]) * mul_dc[c];
          const float prod = ymul * ry[x];
          r[x] = prod + dq;
        }
      }
    }
    DumpImage3F(fdc.dc, stage(PIKGPU_STAGE_DC_DEQUANT), bx, by);
    AdaptiveDCReconstruction(fdc.dc, quantizer, pool.get());
    DumpImage3F(fdc.dc, stage(PIKGPU_STAGE_DC), bx, by);
  }

  // Groups.
  Image3F idct(DivCeil(bx * kBlockDim, downsample),
               DivCeil(by * kBlockDim, downsample));
  Image3F coeffs_dump;
  const bool want_coeffs =
      stage(PIKGPU_STAGE_COEFFS) != nullptr;
  if (want_coeffs) coeffs_dump = Image3F(bx * 64, by);
  const size_t xsize_groups = DivCeil<size_t>(xsize, kGroupDim);
  const size_t ysize_groups = DivCeil<size_t>(ysize, kGroupDim);
  const size_t num_groups = xsize_groups * ysize_groups;
  std::vector<GroupDecCache> caches(std::max<size_t>(1, NumThreads(pool.get())));
  RunOnPool(
      pool.get(), 0, num_groups,
      [&](const int group_index, const int thread) {
        const size_t gx = group_index % xsize_groups;
        const size_t gy = group_index / xsize_groups;
        const Rect rect(gx * kGroupDim, gy * kGroupDim, kGroupDim, kGroupDim,
                        xsize, ysize);
        const Rect padded_rect(rect.x0(), rect.y0(),
                               DivCeil(rect.xsize(), kBlockDim) * kBlockDim,
                               DivCeil(rect.ysize(), kBlockDim) * kBlockDim);
        const Rect block_group_rect(
            padded_rect.x0() / kBlockDim, padded_rect.y0() / kBlockDim,
            padded_rect.xsize() / kBlockDim, padded_rect.ysize() / kBlockDim);
        const Rect group_in_color_tiles(
            block_group_rect.x0() / kColorTileDimInBlocks,
            block_group_rect.y0() / kColorTileDimInBlocks,
            DivCeil(block_group_rect.xsize(), kColorTileDimInBlocks),
            DivCeil(block_group_rect.ysize(), kColorTileDimInBlocks));
        GroupDecCache* gdc = &caches[thread];
        Quantizer quant = quantizer.Copy(block_group_rect);
        Image3S group_ac(block_group_rect.xsize() * 64,
                         block_group_rect.ysize());
        for (int c = 0; c < 3; ++c) {
          for (size_t y = 0; y < block_group_rect.ysize(); ++y) {
            memcpy(group_ac.PlaneRow(c, y),
                   f->ac[c] + (block_group_rect.y0() + y) * f->ac_stride +
                       block_group_rect.x0() * 64,
                   block_group_rect.xsize() * 64 * sizeof(int16_t));
          }
        }
        InitializeDecCache(fdc, padded_rect, gdc);
        DequantImageAC(quant, cmap, group_in_color_tiles, group_ac, &fdc, gdc,
                       padded_rect, nullptr);
        // Later passes: Dequant<false>::DoAC stores each coefficient's
        // dequantised value of the pass plus the value already there
        // (pik/compressed_image.cc:695-703); DequantImageAC (the first-pass
        // form) computes the pass's values into a second cache.
        for (int p = 0; p < num_extra; ++p) {
          for (int c = 0; c < 3; ++c) {
            for (size_t y = 0; y < block_group_rect.ysize(); ++y) {
              memcpy(group_ac.PlaneRow(c, y),
                     extra_ac[3 * p + c] +
                         (block_group_rect.y0() + y) * extra_stride +
                         block_group_rect.x0() * 64,
                     block_group_rect.xsize() * 64 * sizeof(int16_t));
            }
          }
          GroupDecCache pass_cache;
          InitializeDecCache(fdc, padded_rect, &pass_cache);
          DequantImageAC(quant, cmap, group_in_color_tiles, group_ac, &fdc,
                         &pass_cache, padded_rect, nullptr);
          for (int c = 0; c < 3; ++c) {
            for (size_t y = 0; y < block_group_rect.ysize(); ++y) {
              const float* src = pass_cache.ac.ConstPlaneRow(c, y);
              float* dst = gdc->ac.PlaneRow(c, y);
              for (size_t x = 0; x < block_group_rect.xsize() * 64; ++x) {
                dst[x] = src[x] + dst[x];
              }
            }
          }
        }
        const Rect opsin_rect(padded_rect.x0() / downsample,
                              padded_rect.y0() / downsample,
                              DivCeil(padded_rect.xsize(), downsample),
                              DivCeil(padded_rect.ysize(), downsample));
        ReconOpsinImage(frame_header, group_header, quant, block_group_rect,
                        &fdc, gdc, &idct, opsin_rect, nullptr, downsample);
        if (want_coeffs) {
          for (int c = 0; c < 3; ++c) {
            for (size_t y = 0; y < block_group_rect.ysize(); ++y) {
              memcpy(coeffs_dump.PlaneRow(c, block_group_rect.y0() + y) +
                         block_group_rect.x0() * 64,
                     gdc->ac.ConstPlaneRow(c, y),
                     block_group_rect.xsize() * 64 * sizeof(float));
            }
          }
        }
      },
      "recon groups");
  if (want_coeffs) {
    DumpImage3F(coeffs_dump, stage(PIKGPU_STAGE_COEFFS), bx * 64, by);
  }
  DumpImage3F(idct, stage(PIKGPU_STAGE_IDCT), bx * 8, by * 8);

  if (downsample != 1) {
    // FinalizeFrameDecoding's downsampled branch: the dictionary is empty
    // (checked above), no AR, no Gaborish.
    const size_t ox = DivCeil<size_t>(xsize, downsample);
    const size_t oy = DivCeil<size_t>(ysize, downsample);
    idct.ShrinkTo(ox, oy);
    OpsinToLinear(&idct, pool.get());
    if (IsGray(f)) ForceGray(&idct);
    DumpImage3F(idct, out, ox, oy);
    return 0;
  }

  // FinalizeFrameDecoding steps (pik/compressed_image.cc:1110-1181).
  if (frame_header.have_adaptive_reconstruction) {
    Image3F smoothed;
    const Image3F* guide = &idct;
    if (frame_header.gaborish != GaborishStrength::kOff) {
      if (!ConvolveGaborish(idct, frame_header.gaborish, pool.get(),
                            &smoothed)) {
        return Fail("ConvolveGaborish failed");
      }
      guide = &smoothed;
    }
    // The EPF quantises its guide bytes from `idct` itself when
    // use_sharpened is set (AdaptiveFilter, pik/epf_target.cc:1136-1137).
    DumpImage3F(frame_header.epf_params.use_sharpened ? idct : *guide,
                stage(PIKGPU_STAGE_GUIDE), bx * 8, by * 8);
    Image3F filtered;
    AdaptiveReconstructionAux aux;
    aux.filtered = &filtered;
    Image3F reconstructed = AdaptiveReconstruction(
        *guide, idct, quantizer, fdc.raw_quant_field,
        fdc.dequant_control_field, fdc.dequant_map, fdc.ar_sigma_lut_ids,
        fdc.ac_strategy, frame_header.epf_params, pool.get(), &aux);
    DumpImage3F(filtered, stage(PIKGPU_STAGE_EPF), bx * 8, by * 8);
    idct = std::move(reconstructed);
  }
  DumpImage3F(idct, stage(PIKGPU_STAGE_AR), bx * 8, by * 8);
  const BlockDictionary dictionary = ImportDots(*f);
  dictionary.AddTo(&idct, 1);
  DumpImage3F(idct, stage(PIKGPU_STAGE_DOTS), bx * 8, by * 8);
  if (frame_header.gaborish != GaborishStrength::kOff) {
    Image3F gaborished;
    if (!ConvolveGaborish(idct, frame_header.gaborish, pool.get(),
                          &gaborished)) {
      return Fail("ConvolveGaborish failed");
    }
    idct = std::move(gaborished);
  }
  DumpImage3F(idct, stage(PIKGPU_STAGE_GABORISH), bx * 8, by * 8);
  idct.ShrinkTo(xsize, ysize);
  OpsinToLinear(&idct, pool.get());
  if (IsGray(f)) ForceGray(&idct);
  DumpImage3F(idct, stage(PIKGPU_STAGE_LINEAR), xsize, ysize);
  DumpImage3F(idct, out, xsize, ysize);
  return 0;
}

API int pikref_recon(const pikgpu_frame* f, int num_threads,
                     float* const* stages, float* const* out) {
  return Recon(f, num_threads, 1, stages, out);
}

// With AC passes after the frame's own: extra_ac[3 * p + c] is plane c of
// extra pass p (rows of extra_stride int16 in the frame's AC layout).
API int pikref_recon_passes(const pikgpu_frame* f, int num_threads,
                            int num_extra, const int16_t* const* extra_ac,
                            size_t extra_stride, float* const* stages,
                            float* const* out) {
  if (num_extra < 0 || (num_extra > 0 && extra_ac == nullptr)) {
    return Fail("bad extra passes");
  }
  return Recon(f, num_threads, 1, stages, out, num_extra, extra_ac,
               extra_stride);
}

// Downsampled decode: `out` receives 3 planes of ceil(xsize / downsample) x
// ceil(ysize / downsample).
API int pikref_recon_downsampled(const pikgpu_frame* f, int num_threads,
                                 int downsample, float* const* out) {
  return Recon(f, num_threads, size_t(downsample), nullptr, out);
}

// Reference EPF weight for (sigma, sad) (WeightFast via MulTable,
// pik/epf_target.cc:304-421), for brute-force checks of the restatement.
API float pikref_epf_weight(int sigma, int sad) {
  return Dispatch(TargetBitfield().Best(), EdgePreservingFilterTest(), sigma,
                  sad);
}

// Reference variable transforms: IDCT (dir=0) or forward DCT (dir=1) of one
// block of the given raw strategy, in the reference's coefficient layout
// (coefficient row stride = covered_blocks_x * 64, pixel stride = 8 *
// covered_blocks_x).
API int pikref_transform(int raw_strategy, int dir, const float* in,
                         float* out) {
  AcStrategy acs = AcStrategy::FromRawStrategy(raw_strategy);
  const size_t cbx = acs.covered_blocks_x();
  if (dir == 0) {
    acs.TransformToPixels(in, cbx * 64, out, cbx * 8);
  } else {
    acs.TransformFromPixels(in, cbx * 8, out, cbx * 64);
  }
  return 0;
}

// Reference ConvolveGaborish on one image (3 planes, dense xs*ys).
API int pikref_gaborish(int strength, const float* const* in, size_t xs,
                        size_t ys, float* const* out) {
  Image3F img(xs, ys);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < ys; ++y) {
      memcpy(img.PlaneRow(c, y), in[c] + y * xs, xs * 4);
    }
  }
  Image3F res;
  if (!ConvolveGaborish(img, GaborishStrength(strength), nullptr, &res)) {
    return Fail("gaborish failed");
  }
  DumpImage3F(res, out, xs, ys);
  return 0;
}

// Reference OpsinToLinear in place on dense planes.
API int pikref_opsin_to_linear(float* const* planes, size_t xs, size_t ys) {
  Image3F img(xs, ys);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < ys; ++y) {
      memcpy(img.PlaneRow(c, y), planes[c] + y * xs, xs * 4);
    }
  }
  OpsinToLinear(&img, nullptr);
  DumpImage3F(img, planes, xs, ys);
  return 0;
}

// The encoder's coefficient computation as RoundtripImage runs it
// (pik/adaptive_quantization.cc:580-651): InitializeFrameEncCache on the
// whole XYB image, then per group InitializeEncCache and ComputeCoefficients
// with the group's Quantizer copy and colour-tile rect. `opsin` holds three
// dense planes of the 8-padded image (bx * 8 by by * 8); the frame supplies
// the side information (its ac / dc are ignored). `ac_out` receives the
// quantized AC (three dense planes of by rows of bx * 64), `dc_out` the
// quantized DC (three dense bx by by planes); either may be NULL.
API int pikref_quantize(const pikgpu_frame* f, const float* const* opsin,
                        int16_t* const* ac_out, int16_t* const* dc_out) {
  if (f == nullptr || opsin == nullptr) return Fail("null argument");
  if (f->num_dots != 0 || IsGray(f)) {
    return Fail("quantize: frames without dots, not grayscale");
  }
  InitEpfOnce();
  const size_t xsize = f->xsize, ysize = f->ysize;
  const size_t bx = DivCeil<size_t>(xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(ysize, kBlockDim);
  DequantMatrices matrices(/*need_inv_matrices=*/true);
  {
    std::vector<float> t;
    std::vector<uint32_t> o;
    ExportTables(matrices, &t, &o);
    if (f->dequant_table_size != t.size() ||
        memcmp(f->dequant_table, t.data(), t.size() * 4) != 0 ||
        f->num_dequant_offsets != o.size() ||
        memcmp(f->dequant_offsets, o.data(), o.size() * 4) != 0) {
      return Fail("frame does not use the default dequant tables");
    }
  }
  FrameHeader frame_header;
  frame_header.flags = 0;
  frame_header.predict_lf = f->predict_lf != 0;
  frame_header.predict_hf = f->predict_hf != 0;
  GroupHeader group_header;
  Quantizer quantizer =
      DecoderQuantizer(&matrices, bx, by, f->quant_dc, f->global_scale);
  {
    ImageI qf(bx, by);
    for (size_t y = 0; y < by; ++y) {
      memcpy(qf.Row(y), f->quant_field + y * f->quant_field_stride, bx * 4);
    }
    quantizer.SetRawQuantField(std::move(qf));
  }
  ColorCorrelationMap cmap(xsize, ysize);
  for (size_t y = 0; y < cmap.ytox_map.ysize(); ++y) {
    memcpy(cmap.ytox_map.Row(y), f->ytox_map + y * f->ytox_stride,
           cmap.ytox_map.xsize() * 4);
    memcpy(cmap.ytob_map.Row(y), f->ytob_map + y * f->ytob_stride,
           cmap.ytob_map.xsize() * 4);
  }
  cmap.ytox_dc = f->ytox_dc;
  cmap.ytob_dc = f->ytob_dc;
  AcStrategyImage ac_strategy;
  {
    ImageB raw(bx, by);
    for (size_t y = 0; y < by; ++y) {
      for (size_t x = 0; x < bx; ++x) {
        raw.Row(y)[x] = f->ac_strategy[y * f->ac_strategy_stride + x] >> 4;
      }
    }
    ac_strategy = AcStrategyImage(bx, by);
    ac_strategy.SetFromRaw(Rect(0, 0, bx, by), raw);
  }
  const size_t tx = DivCeil<size_t>(xsize, kColorTileDim);
  const size_t ty = DivCeil<size_t>(ysize, kColorTileDim);
  FrameEncCache fec;
  fec.dequant_control_field = ImageB(tx, ty);
  for (size_t y = 0; y < ty; ++y) {
    memcpy(fec.dequant_control_field.Row(y),
           f->dequant_cf + y * f->dequant_cf_stride, tx);
  }
  memcpy(fec.dequant_map, f->dequant_map, sizeof(fec.dequant_map));
  Image3F img(bx * kBlockDim, by * kBlockDim);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < by * kBlockDim; ++y) {
      memcpy(img.PlaneRow(c, y), opsin[c] + y * bx * kBlockDim,
             bx * kBlockDim * 4);
    }
  }
  BlockDictionary dictionary;
  InitializeFrameEncCache(frame_header, img, ac_strategy, quantizer, cmap,
                          dictionary, nullptr, &fec, nullptr);
  const size_t xsize_groups = DivCeil<size_t>(xsize, kGroupDim);
  const size_t ysize_groups = DivCeil<size_t>(ysize, kGroupDim);
  for (size_t gy = 0; gy < ysize_groups; ++gy) {
    for (size_t gx = 0; gx < xsize_groups; ++gx) {
      const Rect group_rect(gx * kGroupDim, gy * kGroupDim, kGroupDim,
                            kGroupDim, bx * kBlockDim, by * kBlockDim);
      const Rect block_rect(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                            kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
      EncCache cache;
      InitializeEncCache(frame_header, group_header, fec, group_rect, &cache);
      Quantizer quant = quantizer.Copy(block_rect);
      const Rect tiles(block_rect.x0() / kColorTileDimInBlocks,
                       block_rect.y0() / kColorTileDimInBlocks,
                       DivCeil(block_rect.xsize(), kColorTileDimInBlocks),
                       DivCeil(block_rect.ysize(), kColorTileDimInBlocks));
      ComputeCoefficients(quant, cmap, tiles, nullptr, fec, &cache, nullptr);
      if (ac_out == nullptr) continue;
      for (int c = 0; c < 3; ++c) {
        if (ac_out[c] == nullptr) continue;
        for (size_t y = 0; y < block_rect.ysize(); ++y) {
          memcpy(ac_out[c] + (block_rect.y0() + y) * bx * 64 +
                     block_rect.x0() * 64,
                 cache.ac.ConstPlaneRow(c, y), block_rect.xsize() * 64 * 2);
        }
      }
    }
  }
  if (dc_out != nullptr) {
    for (int c = 0; c < 3; ++c) {
      if (dc_out[c] == nullptr) continue;
      for (size_t y = 0; y < by; ++y) {
        memcpy(dc_out[c] + y * bx, fec.dc.ConstPlaneRow(c, y), bx * 2);
      }
    }
  }
  return 0;
}

// Reference LinearToXyb for one pixel (for generator tests).
API void pikref_linear_to_xyb(float r, float g, float b, float* xyb) {
  LinearToXyb(r, g, b, &xyb[0], &xyb[1], &xyb[2]);
}

// ---------------------------------------------------------------------------
// Quant-field search: the judge (Butteraugli), the byte count and the
// reference's own search loop as the comparator of the GPU gradient search.

namespace {

// The reference objects of a frame's side information (its AC / DC / quant
// field are not used).
struct Side {
  std::unique_ptr<DequantMatrices> matrices;
  FrameHeader header;
  GroupHeader group_header;
  std::unique_ptr<ColorCorrelationMap> cmap;
  AcStrategyImage acs;
  ImageB luts;
  ImageB dcf;
  uint8_t dqmap[kMaxQuantControlFieldValue][256];
};

int ImportSide(const pikgpu_frame* f, Side* s) {
  if (f == nullptr) return Fail("null frame");
  if (f->num_dots != 0 || IsGray(f)) {
    return Fail("search: frames without dots, not grayscale");
  }
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  s->matrices.reset(new DequantMatrices(/*need_inv_matrices=*/true));
  std::vector<float> t;
  std::vector<uint32_t> o;
  ExportTables(*s->matrices, &t, &o);
  if (f->dequant_table_size != t.size() ||
      memcmp(f->dequant_table, t.data(), t.size() * 4) != 0 ||
      f->num_dequant_offsets != o.size() ||
      memcmp(f->dequant_offsets, o.data(), o.size() * 4) != 0) {
    return Fail("frame does not use the default dequant tables");
  }
  s->header.flags = 0;
  s->header.gaborish = GaborishStrength(f->gaborish);
  s->header.predict_lf = f->predict_lf != 0;
  s->header.predict_hf = f->predict_hf != 0;
  s->header.have_adaptive_reconstruction = f->adaptive_reconstruction != 0;
  s->header.epf_params.enable_adaptive = f->epf_enable_adaptive != 0;
  s->header.epf_params.use_sharpened = f->epf_use_sharpened != 0;
  s->cmap.reset(new ColorCorrelationMap(f->xsize, f->ysize));
  for (size_t y = 0; y < s->cmap->ytox_map.ysize(); ++y) {
    memcpy(s->cmap->ytox_map.Row(y), f->ytox_map + y * f->ytox_stride,
           s->cmap->ytox_map.xsize() * 4);
    memcpy(s->cmap->ytob_map.Row(y), f->ytob_map + y * f->ytob_stride,
           s->cmap->ytob_map.xsize() * 4);
  }
  s->cmap->ytox_dc = f->ytox_dc;
  s->cmap->ytob_dc = f->ytob_dc;
  ImageB raw(bx, by);
  s->luts = ImageB(bx, by);
  for (size_t y = 0; y < by; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      raw.Row(y)[x] = f->ac_strategy[y * f->ac_strategy_stride + x] >> 4;
      s->luts.Row(y)[x] = f->sigma_lut_ids[y * f->sigma_lut_stride + x];
    }
  }
  s->acs = AcStrategyImage(bx, by);
  s->acs.SetFromRaw(Rect(0, 0, bx, by), raw);
  const size_t tx = DivCeil<size_t>(f->xsize, kColorTileDim);
  const size_t ty = DivCeil<size_t>(f->ysize, kColorTileDim);
  s->dcf = ImageB(tx, ty);
  for (size_t y = 0; y < ty; ++y) {
    memcpy(s->dcf.Row(y), f->dequant_cf + y * f->dequant_cf_stride, tx);
  }
  memcpy(s->dqmap, f->dequant_map, sizeof(s->dqmap));
  return 0;
}

Image3F ImportPlanes(const float* const* p, size_t xs, size_t ys) {
  Image3F img(xs, ys);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < ys; ++y) {
      memcpy(img.PlaneRow(c, y), p[c] + y * xs, xs * 4);
    }
  }
  return img;
}

CompressParams SearchParams(float target) {
  CompressParams cp;
  cp.butteraugli_distance = target;
  return cp;
}

}  // namespace

// Butteraugli distance of `linear` (dense xs x ys planes, linear RGB in
// [0, 255]) from the image whose XYB planes are `opsin` (dense, unpadded),
// as the reference encoder's comparator measures it (ButteraugliComparator,
// pik/butteraugli_comparator.cc:23-44, hf_asymmetry 1 and the default
// intensity multiplier, pik/pik_params.h:116-136). `distmap` (dense xs x ys,
// may be NULL) receives the per-pixel map.
API int pikref_butteraugli(const float* const* opsin, const float* const* linear,
                           size_t xs, size_t ys, float* distmap,
                           double* distance) {
  if (opsin == nullptr || linear == nullptr || distance == nullptr) {
    return Fail("null argument");
  }
  const CompressParams cp;
  ButteraugliComparator comparator(ImportPlanes(opsin, xs, ys),
                                   cp.hf_asymmetry,
                                   cp.GetIntensityMultiplier());
  comparator.Compare(ImportPlanes(linear, xs, ys));
  *distance = comparator.distance();
  if (distmap != nullptr) {
    for (size_t y = 0; y < ys; ++y) {
      memcpy(distmap + y * xs, comparator.distmap().ConstRow(y), xs * 4);
    }
  }
  return 0;
}

// The reference's initial quant field (InitialQuantField,
// pik/adaptive_quantization.cc:1047-1060) of the unpadded XYB image `opsin`
// (dense xs x ys planes) at Butteraugli distance `target`: by x bx floats.
API int pikref_initial_quant_field(const float* const* opsin, size_t xs,
                                   size_t ys, float target, float* qf) {
  if (opsin == nullptr || qf == nullptr) return Fail("null argument");
  const CompressParams cp = SearchParams(target);
  const ImageF field =
      InitialQuantField(target, cp.GetIntensityMultiplier(),
                        ImportPlanes(opsin, xs, ys), cp, nullptr, 1.0);
  for (size_t y = 0; y < field.ysize(); ++y) {
    memcpy(qf + y * field.xsize(), field.ConstRow(y), field.xsize() * 4);
  }
  return 0;
}

// InitialQuantDC (pik/adaptive_quantization.cc:1040-1045) at `target`.
API float pikref_initial_quant_dc(float target) {
  return InitialQuantDC(target, SearchParams(target).GetIntensityMultiplier());
}

// The reference's colour correlation map of the padded XYB image
// (FindBestColorCorrelationMap, pik/color_correlation.h:101): ytox / ytob
// per 64x64 tile (ty x tx each) and the DC factors.
API int pikref_color_correlation(const float* const* opsin, size_t xs,
                                 size_t ys, int32_t* ytox, int32_t* ytob,
                                 int32_t* dc) {
  if (opsin == nullptr || ytox == nullptr || ytob == nullptr || dc == nullptr) {
    return Fail("null argument");
  }
  DequantMatrices matrices(/*need_inv_matrices=*/true);
  ColorCorrelationMap cmap(xs, ys);
  FindBestColorCorrelationMap(ImportPlanes(opsin, xs, ys), matrices, &cmap);
  const size_t tx = cmap.ytox_map.xsize(), ty = cmap.ytox_map.ysize();
  for (size_t y = 0; y < ty; ++y) {
    memcpy(ytox + y * tx, cmap.ytox_map.ConstRow(y), tx * 4);
    memcpy(ytob + y * tx, cmap.ytob_map.ConstRow(y), tx * 4);
  }
  dc[0] = cmap.ytox_dc;
  dc[1] = cmap.ytob_dc;
  return 0;
}

// Encoded size of a frame's quantized data with the reference's entropy
// coder: *ac_bytes sums EncodeToBitstream over the 512-px groups (coefficient
// orders, histograms and AC tokens, pik/compressed_image.cc:457-573, as
// PixelsToPikGroup encodes each group, pik/pik_frame.cc:345-411);
// *dc_bytes the DC groups (quantized DC, AC strategy, quant field and sigma
// LUT ids as CompressDCGroup codes them, pik/compressed_dc.cc:233-273, group
// TOC excluded) plus the quantizer's three bytes and the colour correlation
// map (pik/pik_frame.cc:560-572).
API int pikref_encoded_bytes(const pikgpu_frame* f, size_t* ac_bytes,
                             size_t* dc_bytes) {
  Side side;
  if (ImportSide(f, &side) != 0) return -1;
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  Quantizer quantizer = DecoderQuantizer(side.matrices.get(), bx, by,
                                         f->quant_dc, f->global_scale);
  {
    ImageI qf(bx, by);
    for (size_t y = 0; y < by; ++y) {
      memcpy(qf.Row(y), f->quant_field + y * f->quant_field_stride, bx * 4);
    }
    quantizer.SetRawQuantField(std::move(qf));
  }
  size_t ac = 0;
  const size_t gxs = DivCeil<size_t>(f->xsize, kGroupDim);
  const size_t gys = DivCeil<size_t>(f->ysize, kGroupDim);
  for (size_t gy = 0; gy < gys; ++gy) {
    for (size_t gx = 0; gx < gxs; ++gx) {
      const Rect group(gx * kGroupDim, gy * kGroupDim, kGroupDim, kGroupDim,
                       bx * kBlockDim, by * kBlockDim);
      const Rect blocks(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                        kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
      EncCache cache;
      cache.xsize_blocks = blocks.xsize();
      cache.ysize_blocks = blocks.ysize();
      cache.ac_strategy = side.acs.Copy(blocks);
      cache.ac = Image3S(blocks.xsize() * 64, blocks.ysize());
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < blocks.ysize(); ++y) {
          memcpy(cache.ac.PlaneRow(c, y),
                 f->ac[c] + (blocks.y0() + y) * f->ac_stride + blocks.x0() * 64,
                 blocks.xsize() * 64 * 2);
        }
      }
      const PaddedBytes bytes = EncodeToBitstream(
          cache, Rect(0, 0, group.xsize(), group.ysize()),
          quantizer.Copy(blocks), NoiseParams(), /*fast_mode=*/false,
          /*handler=*/nullptr);
      ac += bytes.size();
    }
  }
  // DC groups as CompressDCGroup codes them without the lossless DC coder
  // (use_new_dc off, the default, pik/compressed_dc.cc:233-273; the
  // single-image manager gives no hints): DC residuals, then the control
  // fields (strategy, quant field, sigma LUT ids) with their histograms.
  Image3S dc(bx, by);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < by; ++y) {
      memcpy(dc.PlaneRow(c, y), f->dc[c] + y * f->dc_stride, bx * 2);
    }
  }
  size_t dcb = 0;
  const size_t dgx = DivCeil<size_t>(bx, kDcGroupDimInBlocks);
  const size_t dgy = DivCeil<size_t>(by, kDcGroupDimInBlocks);
  for (size_t gy = 0; gy < dgy; ++gy) {
    for (size_t gx = 0; gx < dgx; ++gx) {
      const Rect rect(gx * kDcGroupDimInBlocks, gy * kDcGroupDimInBlocks,
                      kDcGroupDimInBlocks, kDcGroupDimInBlocks, bx, by);
      Image3S residuals(rect.xsize(), rect.ysize());
      ShrinkDC(rect, dc, &residuals);
      const std::string dc_code =
          EncodeImageData(Rect(residuals), residuals, nullptr);
      std::vector<std::vector<Token>> tokens(1);
      TokenizeAcStrategy(rect, side.acs, nullptr, &tokens[0]);
      TokenizeQuantField(rect, quantizer.RawQuantField(), nullptr, side.acs,
                         &tokens[0]);
      TokenizeARParameters(rect, side.luts, side.acs, &tokens[0]);
      std::vector<uint8_t> context_map;
      std::vector<ANSEncodingData> codes;
      const std::string histo = BuildAndEncodeHistograms(
          kNumControlFieldContexts, tokens, &codes, &context_map, nullptr);
      const std::string fields =
          WriteTokens(tokens[0], codes, context_map, nullptr);
      dcb += dc_code.size() + histo.size() + fields.size();
    }
  }
  const std::string qcode = quantizer.Encode(nullptr);
  const std::string cmap_code =
      EncodeColorMap(side.cmap->ytob_map, Rect(side.cmap->ytob_map),
                     side.cmap->ytob_dc, nullptr) +
      EncodeColorMap(side.cmap->ytox_map, Rect(side.cmap->ytox_map),
                     side.cmap->ytox_dc, nullptr);
  if (ac_bytes) *ac_bytes = ac;
  if (dc_bytes) *dc_bytes = dcb + qcode.size() + cmap_code.size();
  return 0;
}

namespace {

// TileDistMap (pik/adaptive_quantization.cc:420-482; in an unnamed
// namespace there): per transform, the 16-norm of the distance map over its
// pixels widened by `margin` (edge pixels weighted 0.98, corners 0.7 when
// margin != 0), times 1.2, written to every block the transform covers.
ImageF TileDistances(const ImageF& dist, int margin,
                     const AcStrategyImage& acs) {
  const int tx = int(DivCeil<size_t>(dist.xsize(), 8));
  const int ty = int(DivCeil<size_t>(dist.ysize(), 8));
  ImageF out(tx, ty);
  for (int by = 0; by < ty; ++by) {
    for (int bx = 0; bx < tx; ++bx) {
      const AcStrategy a = acs.ConstRow(by)[bx];
      if (!a.IsFirstBlock()) continue;
      const int y0 = std::max(0, 8 * by - margin);
      const int y1 = std::min<int>(dist.ysize(),
                                   8 * (by + int(a.covered_blocks_y())) + margin);
      const int x0 = std::max(0, 8 * bx - margin);
      const int x1 = std::min<int>(dist.xsize(),
                                   8 * (bx + int(a.covered_blocks_x())) + margin);
      float norm = 0.0f;
      double weight = 0.0;
      for (int y = y0; y < y1; ++y) {
        const bool yedge = margin != 0 && (y == y0 || y == y1 - 1);
        const float* row = dist.ConstRow(y);
        for (int x = x0; x < x1; ++x) {
          const bool xedge = margin != 0 && (x == x0 || x == x1 - 1);
          const float w = yedge ? (xedge ? 0.7f : 0.98f) : (xedge ? 0.98f : 1.0f);
          float v = row[x];
          v *= v;
          v *= v;
          v *= v;
          v *= v;
          norm += w * v;
          weight += w;
        }
      }
      if (weight == 0) weight = 1;
      const double d = 1.2 * std::pow(norm / weight, 1.0 / 16);
      for (size_t iy = 0; iy < a.covered_blocks_y(); ++iy) {
        for (size_t ix = 0; ix < a.covered_blocks_x(); ++ix) {
          out.Row(by + iy)[bx + ix] = d;
        }
      }
    }
  }
  return out;
}

// AdjustQuantField (pik/adaptive_quantization.cc:554-578): a multi-block
// transform takes the maximum of the values over its blocks.
void MaxOverTransforms(const AcStrategyImage& acs, ImageF* field) {
  for (size_t y = 0; y < field->ysize(); ++y) {
    for (size_t x = 0; x < field->xsize(); ++x) {
      const AcStrategy a = acs.ConstRow(y)[x];
      if (!a.IsFirstBlock()) continue;
      float m = field->Row(y)[x];
      for (size_t iy = 0; iy < a.covered_blocks_y(); ++iy) {
        for (size_t ix = 0; ix < a.covered_blocks_x(); ++ix) {
          m = std::max(m, field->Row(y + iy)[x + ix]);
        }
      }
      for (size_t iy = 0; iy < a.covered_blocks_y(); ++iy) {
        for (size_t ix = 0; ix < a.covered_blocks_x(); ++ix) {
          field->Row(y + iy)[x + ix] = m;
        }
      }
    }
  }
}

}  // namespace

namespace {

// One round trip of the search: the frame with `quantizer`'s serialized
// scalars and field, its coefficients quantized from `opsin` by the
// reference encoder (pikref_quantize), then decoded by the reference
// (Recon) into `linear`.
int Roundtrip(const pikgpu_frame* f, const Quantizer& quantizer,
              const float* const* opsin, int num_threads,
              std::vector<int16_t>* ac, std::vector<int16_t>* dc,
              Image3F* linear) {
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  pikgpu_frame g = *f;
  const std::string b = quantizer.Encode(nullptr);
  g.global_scale = ((uint8_t(b[0]) << 8) | uint8_t(b[1])) + 1;
  g.quant_dc = uint8_t(b[2]) + 1;
  g.quant_field = quantizer.RawQuantField().ConstRow(0);
  g.quant_field_stride = quantizer.RawQuantField().PixelsPerRow();
  ac->assign(3 * by * bx * 64, 0);
  dc->assign(3 * by * bx, 0);
  int16_t* acp[3];
  int16_t* dcp[3];
  for (int c = 0; c < 3; ++c) {
    acp[c] = ac->data() + c * by * bx * 64;
    dcp[c] = dc->data() + c * by * bx;
    g.ac[c] = acp[c];
    g.dc[c] = dcp[c];
  }
  g.ac_stride = bx * 64;
  g.dc_stride = bx;
  if (pikref_quantize(&g, opsin, acp, dcp) != 0) return -1;
  float* out[3];
  std::vector<float> buf(3 * f->xsize * f->ysize);
  for (int c = 0; c < 3; ++c) out[c] = buf.data() + c * f->xsize * f->ysize;
  if (Recon(&g, num_threads, 1, nullptr, out) != 0) return -1;
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < f->ysize; ++y) {
      memcpy(linear->PlaneRow(c, y), out[c] + y * f->xsize, f->xsize * 4);
    }
  }
  return 0;
}

}  // namespace

// The reference encoder's quant-field search (FindBestQuantization,
// pik/adaptive_quantization.cc:677-853), restated over the reference's own
// functions because the function itself cannot be linked here (its debug
// heat maps reach the colour-management code). Each round sets the field
// through Quantizer::SetQuantField, round-trips the image through the
// reference's RoundtripImage, compares with the reference Butteraugli, tiles
// the distance map (TileDistMap) and rescales every block's quant value by
// its tile distance relative to the target: upward where the distance is
// above the target, downward by a round-dependent power of the ratio below
// it, within bounds derived from the initial field; from round 6 on, an
// increase that made the nearby distance worse by more than 1.5 % is taken
// back; at round 5 the field is pulled 60 % of the way back toward the
// initial one where it dropped below that.
//
// Inputs: the frame's side information (strategy, sigma LUTs, colour
// correlation, dequant control field / map, header flags); `opsin` the
// padded XYB image the encoder quantizes, `opsin_orig` the unpadded one the
// comparator holds (dense planes); `init` the initial float field (by x bx).
// Outputs: the final integer field (qf_out, by x bx), global scale and
// DC quant (scalars[0..1]), the last measured distance and the loop's wall
// time (excluding set-up). `rounds` Butteraugli rounds (the reference runs
// max_butteraugli_iters + 1 = 12).
API int pikref_find_best_quantization(const pikgpu_frame* f,
                                      const float* const* opsin,
                                      const float* const* opsin_orig,
                                      const float* init, float target,
                                      int rounds, int num_threads,
                                      int32_t* qf_out, int32_t* scalars,
                                      double* distance, double* seconds) {
  if (opsin == nullptr || opsin_orig == nullptr || init == nullptr ||
      qf_out == nullptr || scalars == nullptr) {
    return Fail("null argument");
  }
  if (rounds < 1) return Fail("rounds < 1");
  InitEpfOnce();
  Side side;
  if (ImportSide(f, &side) != 0) return -1;
  const size_t xs = f->xsize, ys = f->ysize;
  const size_t bx = DivCeil<size_t>(xs, kBlockDim);
  const size_t by = DivCeil<size_t>(ys, kBlockDim);
  const Image3F orig = ImportPlanes(opsin_orig, xs, ys);
  const Image3F padded = ImportPlanes(opsin, bx * kBlockDim, by * kBlockDim);
  std::unique_ptr<ThreadPool> pool = MakePool(num_threads);
  const CompressParams cp = SearchParams(target);
  Quantizer quantizer(side.matrices.get(), bx, by);
  std::vector<int16_t> q_ac, q_dc;
  Image3F linear(xs, ys);

  const auto t0 = std::chrono::steady_clock::now();
  ButteraugliComparator comparator(orig, cp.hf_asymmetry,
                                   cp.GetIntensityMultiplier());
  const float quant_dc = InitialQuantDC(target, cp.GetIntensityMultiplier());
  ImageF field(bx, by);
  for (size_t y = 0; y < by; ++y) memcpy(field.Row(y), init + y * bx, bx * 4);
  MaxOverTransforms(side.acs, &field);
  const ImageF first = CopyImage(field);
  float lo0, hi0;
  ImageMinMax(first, &lo0, &hi0);
  const float dev = std::sqrt(250.0f / (hi0 / lo0));
  const float asym = std::min(2.0f, dev);
  const float qf_lo = lo0 / (asym * dev), qf_hi = hi0 * (dev / asym);
  ImageF tiles, local, prev_field = CopyImage(first), prev_local;
  const int kMargin[12] = {0, 0, 0, 1, 2, 1, 1, 1, 0, 0, 0, 0};
  const double kPowBase[3] = {0.97524596113492301, 1.0424361904568509,
                              0.64984804448911193};
  const double kPowSlope[4] = {0.011236980155043978, 0.0061256294105472651,
                               -0.0030115055086858242, 0.06929488142351059};
  double last_distance = 0.0;
  const int last = rounds - 1;
  for (int r = 0; r <= last; ++r) {
    if (quantizer.SetQuantField(quant_dc, QuantField(field))) {
      // RoundtripImage (pik/adaptive_quantization.cc:580-671, unnamed
      // namespace): the frame quantized with this field, then decoded
      if (Roundtrip(f, quantizer, opsin, num_threads, &q_ac, &q_dc, &linear)) {
        return -1;
      }
      comparator.Compare(linear);
      last_distance = comparator.distance();
      tiles = TileDistances(comparator.distmap(), kMargin[std::min(r, 11)],
                            side.acs);
      local = TileDistances(comparator.distmap(), 2, side.acs);
    }
    if (r > 5) {
      for (size_t y = 0; y < by; ++y) {
        float* q = field.Row(y);
        for (size_t x = 0; x < bx; ++x) {
          if (q[x] > prev_field.Row(y)[x] &&
              local.Row(y)[x] > 1.015f * prev_local.Row(y)[x]) {
            q[x] = prev_field.Row(y)[x];
          }
        }
      }
    }
    prev_field = CopyImage(field);
    prev_local = CopyImage(local);
    if (r == last) break;
    if (r == 5) {
      for (size_t y = 0; y < by; ++y) {
        float* q = field.Row(y);
        for (size_t x = 0; x < bx; ++x) {
          const double floor_v = 0.4 * q[x] + 0.6 * first.Row(y)[x];
          if (q[x] < floor_v) q[x] = std::min(std::max(float(floor_v), qf_lo), qf_hi);
        }
      }
    }
    double p = 0.0;
    if (r < 7) {
      p = (r < 3 ? kPowBase[r] : 0.0) +
          (target - 1.0) * (r < 4 ? kPowSlope[r] : 0.0);
      if (p < 0) p = 0;
    }
    for (size_t y = 0; y < by; ++y) {
      float* q = field.Row(y);
      for (size_t x = 0; x < bx; ++x) {
        const float ratio = tiles.Row(y)[x] / target;
        if (ratio >= 1.0f) {
          q[x] *= ratio;
        } else if (p != 0.0) {
          q[x] *= std::pow(ratio, p);
        }
        q[x] = std::min(std::max(q[x], qf_lo), qf_hi);
      }
    }
  }
  quantizer.SetQuantField(quant_dc, QuantField(field));
  const auto t1 = std::chrono::steady_clock::now();
  for (size_t y = 0; y < by; ++y) {
    memcpy(qf_out + y * bx, quantizer.RawQuantField().ConstRow(y), bx * 4);
  }
  int gs = 0, qdc = 0;
  {
    // the serialized scalars (Quantizer::Encode, pik/quantizer.cc:51-61)
    const std::string b = quantizer.Encode(nullptr);
    gs = ((uint8_t(b[0]) << 8) | uint8_t(b[1])) + 1;
    qdc = uint8_t(b[2]) + 1;
  }
  scalars[0] = gs;
  scalars[1] = qdc;
  if (distance) *distance = last_distance;
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

// --------------------------------------------------------------------------
// AC group bitstreams: the reference encoder's EncodeToBitstream per
// 512x512 group (pik/compressed_image.cc:457-573, as EncodePass calls it,
// pik/pik_frame.cc:374-410) and the reference decoder's entropy half of
// DecodeFromBitstream (pik/compressed_image.cc:719-813), for the host
// entropy decoder's parity tests and fixtures.

namespace {

int ImportAcStrategy(const pikgpu_frame* f, AcStrategyImage* acs) {
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  ImageB raw(bx, by);
  for (size_t y = 0; y < by; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      raw.Row(y)[x] = f->ac_strategy[y * f->ac_strategy_stride + x] >> 4;
    }
  }
  *acs = AcStrategyImage(bx, by);
  acs->SetFromRaw(Rect(0, 0, bx, by), raw);
  for (size_t y = 0; y < by; ++y) {
    if (memcmp(acs->ConstRaw().ConstRow(y),
               f->ac_strategy + y * f->ac_strategy_stride, bx) != 0) {
      return Fail("inconsistent ac_strategy bytes");
    }
  }
  return 0;
}

Quantizer FrameQuantizer(const pikgpu_frame* f, const DequantMatrices* m) {
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  Quantizer q = DecoderQuantizer(m, bx, by, f->quant_dc, f->global_scale);
  ImageI qf(bx, by);
  for (size_t y = 0; y < by; ++y) {
    memcpy(qf.Row(y), f->quant_field + y * f->quant_field_stride, bx * 4);
  }
  q.SetRawQuantField(std::move(qf));
  return q;
}

}  // namespace

// Writes the groups' bitstreams back to back into `out` (may be NULL to
// size them) and each group's length into group_size; *total receives the
// sum. `noise` (alpha, gamma, beta; may be NULL) goes into every group's
// noise parameters; fast_mode selects BuildAndEncodeHistogramsFast.
API int pikref_encode_ac_groups(const pikgpu_frame* f, const float* noise,
                                int fast_mode, uint8_t* out, size_t cap,
                                size_t* group_size, size_t* total) {
  if (f == nullptr) return Fail("null frame");
  AcStrategyImage acs;
  if (ImportAcStrategy(f, &acs) != 0) return -1;
  DequantMatrices matrices(/*need_inv_matrices=*/false);
  const Quantizer quantizer = FrameQuantizer(f, &matrices);
  NoiseParams np;
  if (noise != nullptr) {
    np.alpha = noise[0];
    np.gamma = noise[1];
    np.beta = noise[2];
  }
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  const size_t gxs = DivCeil<size_t>(f->xsize, kGroupDim);
  const size_t gys = DivCeil<size_t>(f->ysize, kGroupDim);
  size_t pos = 0;
  for (size_t gy = 0; gy < gys; ++gy) {
    for (size_t gx = 0; gx < gxs; ++gx) {
      const Rect group(gx * kGroupDim, gy * kGroupDim, kGroupDim, kGroupDim,
                       bx * kBlockDim, by * kBlockDim);
      const Rect blocks(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                        kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
      EncCache cache;
      cache.xsize_blocks = blocks.xsize();
      cache.ysize_blocks = blocks.ysize();
      cache.ac_strategy = acs.Copy(blocks);
      cache.ac = Image3S(blocks.xsize() * 64, blocks.ysize());
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < blocks.ysize(); ++y) {
          memcpy(cache.ac.PlaneRow(c, y),
                 f->ac[c] + (blocks.y0() + y) * f->ac_stride + blocks.x0() * 64,
                 blocks.xsize() * 64 * 2);
        }
      }
      const PaddedBytes bytes = EncodeToBitstream(
          cache, Rect(0, 0, group.xsize(), group.ysize()),
          quantizer.Copy(blocks), np, fast_mode != 0, /*handler=*/nullptr);
      if (group_size != nullptr) group_size[gy * gxs + gx] = bytes.size();
      if (out != nullptr) {
        if (pos + bytes.size() > cap) return Fail("output buffer too small");
        memcpy(out + pos, bytes.data(), bytes.size());
      }
      pos += bytes.size();
    }
  }
  if (total != nullptr) *total = pos;
  return 0;
}

namespace {

// One group's pass data as DecodeFromBitstream reads it (DecodeNoise,
// DecodeCoeffOrder x 3, DecodeHistograms, DecodeAC per tile with
// GatherCoefficients, CheckANSFinalState, zero bits to the byte boundary)
// into the group's int16 AC (xsb * 64 x ysb).
int RefDecodeAcGroup(BitReader* br, const AcStrategyImage& acs,
                     const Rect& blocks, Image3S* out, NoiseParams* np) {
  const size_t xsb = blocks.xsize(), ysb = blocks.ysize();
  if (!DecodeNoise(br, np)) return Fail("DecodeNoise failed");
  int order[kOrderContexts * kDCTBlockSize];
  for (size_t c = 0; c < kOrderContexts; ++c) {
    DecodeCoeffOrder(&order[c * kDCTBlockSize], br);
  }
  if (!br->JumpToByteBoundary()) return Fail("order padding");
  ANSCode code;
  std::vector<uint8_t> context_map;
  if (!DecodeHistograms(br, kNumContexts, 256, &code, &context_map)) {
    return Fail("DecodeHistograms failed");
  }
  if (!br->JumpToByteBoundary()) return Fail("histogram padding");
  ANSSymbolReader decoder(&code);
  Image3S tile(kTileDimInBlocks * kDCTBlockSize, kTileDimInBlocks);
  Image3I nz(kTileDimInBlocks, kTileDimInBlocks);
  Image3S& group_ac = *out;
  const size_t stride = tile.PixelsPerRow();
  SIMD_ALIGN int16_t unscattered[AcStrategy::kMaxCoeffArea];
  const size_t ntx = DivCeil(xsb, kTileDimInBlocks);
  const size_t nty = DivCeil(ysb, kTileDimInBlocks);
  for (size_t t = 0; t < ntx * nty; ++t) {
    const Rect rect((t % ntx) * kTileDimInBlocks,
                    (t / ntx) * kTileDimInBlocks, kTileDimInBlocks,
                    kTileDimInBlocks, xsb, ysb);
    if (!DecodeAC(context_map, order, br, &decoder, &tile, rect, &nz)) {
      return Fail("DecodeAC failed");
    }
    for (int c = 0; c < 3; ++c) {
      for (size_t y = 0; y < rect.ysize(); ++y) {
        for (size_t x = 0; x < rect.xsize(); ++x) {
          const AcStrategy s =
              acs.ConstRow(blocks, rect.y0() + y)[rect.x0() + x];
          if (!s.IsFirstBlock()) continue;
          const int16_t* block = tile.ConstPlaneRow(c, y) + x * 64;
          const size_t w = s.covered_blocks_x() * kDCTBlockSize;
          s.GatherCoefficients(block, stride, unscattered, w);
          for (size_t i = 0; i < s.covered_blocks_y(); ++i) {
            memcpy(group_ac.PlaneRow(c, rect.y0() + y + i) +
                       (rect.x0() + x) * 64,
                   unscattered + w * i, w * sizeof(int16_t));
          }
        }
      }
    }
  }
  if (!decoder.CheckANSFinalState()) return Fail("ANS final state");
  if (!br->JumpToByteBoundary()) return Fail("AC padding");
  return 0;
}

}  // namespace

// Decodes the groups' bitstreams (back to back in `data`) with the
// reference's own functions in DecodeFromBitstream's order -- DecodeNoise,
// DecodeCoeffOrder, DecodeHistograms, one ANSSymbolReader through DecodeAC
// of every tile, GatherCoefficients, CheckANSFinalState -- into frame-wide
// int16 AC. With check_dequant set it also runs DecodeFromBitstream<true>
// itself on every group and fails unless its dequantized coefficients equal
// DequantImageAC of the int16 decoded here, bit for bit.
API int pikref_decode_ac_groups(const pikgpu_frame* f, const uint8_t* data,
                                const size_t* group_size,
                                int16_t* const* ac_out, size_t ac_stride,
                                float* noise_out, size_t* consumed,
                                int check_dequant) {
  if (f == nullptr) return Fail("null frame");
  AcStrategyImage acs;
  if (ImportAcStrategy(f, &acs) != 0) return -1;
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  const size_t gxs = DivCeil<size_t>(f->xsize, kGroupDim);
  const size_t gys = DivCeil<size_t>(f->ysize, kGroupDim);

  std::unique_ptr<DequantMatrices> matrices;
  std::unique_ptr<Quantizer> quantizer;
  std::unique_ptr<ColorCorrelationMap> cmap;
  FrameDecCache fdc;
  FrameHeader frame_header;
  GroupHeader group_header;
  if (check_dequant) {
    matrices.reset(new DequantMatrices(/*need_inv_matrices=*/false));
    quantizer.reset(new Quantizer(FrameQuantizer(f, matrices.get())));
    cmap.reset(new ColorCorrelationMap(f->xsize, f->ysize));
    for (size_t y = 0; y < cmap->ytox_map.ysize(); ++y) {
      memcpy(cmap->ytox_map.Row(y), f->ytox_map + y * f->ytox_stride,
             cmap->ytox_map.xsize() * 4);
      memcpy(cmap->ytob_map.Row(y), f->ytob_map + y * f->ytob_stride,
             cmap->ytob_map.xsize() * 4);
    }
    cmap->ytox_dc = f->ytox_dc;
    cmap->ytob_dc = f->ytob_dc;
    fdc.ac_strategy = acs.Copy();
    fdc.raw_quant_field = CopyImage(quantizer->RawQuantField());
    const size_t tx = DivCeil<size_t>(f->xsize, kColorTileDim);
    const size_t ty = DivCeil<size_t>(f->ysize, kColorTileDim);
    fdc.dequant_control_field = ImageB(tx, ty);
    for (size_t y = 0; y < ty; ++y) {
      memcpy(fdc.dequant_control_field.Row(y),
             f->dequant_cf + y * f->dequant_cf_stride, tx);
    }
    memcpy(fdc.dequant_map, f->dequant_map, sizeof(fdc.dequant_map));
  }

  size_t pos = 0;
  for (size_t gy = 0; gy < gys; ++gy) {
    for (size_t gx = 0; gx < gxs; ++gx) {
      const size_t gi = gy * gxs + gx;
      const uint8_t* bytes = data + pos;
      const size_t size = group_size[gi];
      pos += size;
      const Rect blocks(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                        kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
      const size_t xsb = blocks.xsize(), ysb = blocks.ysize();
      BitReader br(bytes, size);
      NoiseParams np;
      Image3S group_ac(xsb * kDCTBlockSize, ysb);
      if (RefDecodeAcGroup(&br, acs, blocks, &group_ac, &np) != 0) return -1;
      if (consumed != nullptr) consumed[gi] = br.Position();
      if (noise_out != nullptr) {
        noise_out[3 * gi + 0] = np.alpha;
        noise_out[3 * gi + 1] = np.gamma;
        noise_out[3 * gi + 2] = np.beta;
      }
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < ysb; ++y) {
          memcpy(ac_out[c] + (blocks.y0() + y) * ac_stride + blocks.x0() * 64,
                 group_ac.ConstPlaneRow(c, y), xsb * 64 * sizeof(int16_t));
        }
      }
      if (!check_dequant) continue;
      const Rect padded_rect(blocks.x0() * kBlockDim, blocks.y0() * kBlockDim,
                             xsb * kBlockDim, ysb * kBlockDim);
      const Rect group_in_color_tiles(
          blocks.x0() / kColorTileDimInBlocks,
          blocks.y0() / kColorTileDimInBlocks,
          DivCeil(xsb, kColorTileDimInBlocks),
          DivCeil(ysb, kColorTileDimInBlocks));
      const Quantizer quant = quantizer->Copy(blocks);
      PaddedBytes compressed(size);
      if (size != 0) memcpy(compressed.data(), bytes, size);
      BitReader br2(compressed.data(), compressed.size());
      GroupDecCache full;
      NoiseParams np2;
      if (!DecodeFromBitstream<true>(frame_header, group_header, compressed,
                                     &br2, padded_rect, /*handler=*/nullptr,
                                     xsb, ysb, *cmap, group_in_color_tiles,
                                     &np2, quant, &fdc, &full, nullptr)) {
        return Fail("DecodeFromBitstream<true> failed");
      }
      GroupDecCache mine;
      mine.InitOnce(xsb, ysb);
      DequantImageAC(quant, *cmap, group_in_color_tiles, group_ac, &fdc,
                     &mine, padded_rect, nullptr);
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < ysb; ++y) {
          if (memcmp(full.ac.ConstPlaneRow(c, y), mine.ac.ConstPlaneRow(c, y),
                     xsb * 64 * sizeof(float)) != 0) {
            return Fail("DecodeFromBitstream<true> disagrees with the "
                        "decoded AC");
          }
        }
      }
      if (np2.alpha != np.alpha || np2.gamma != np.gamma ||
          np2.beta != np.beta) {
        return Fail("DecodeFromBitstream<true> noise parameters differ");
      }
    }
  }
  return 0;
}

// --------------------------------------------------------------------------
// DC groups section: EncodeDCGroups (pik/compressed_dc.cc:352-433) with
// CompressDCGroup (:233-273) for the default old DC coder and no multi-pass
// hints, and the decoding side of DecodeDCGroups (:436-507) with
// DecodeDCGroup (:276-345) up to DequantDC. CompressDCGroup and
// DecodeDCGroup sit in an unnamed namespace there and the public drivers
// need a MultipassManager whose implementations pull in colour management,
// so both are restated here over the reference's own public pieces
// (ShrinkDC / ExpandDC, EncodeImageData / DecodeImage, the control-field
// tokenizers and decoders, BuildAndEncodeHistograms / DecodeHistograms,
// WriteTokens / ANSSymbolReader, SizeCoderT's U32Coder).

namespace {
constexpr uint32_t kDcGroupSizeDistribution = 0x150F0E0C;
}  // namespace

namespace {

// The section (size table + groups) of the frame's quantized DC, strategy,
// quant field and sigma LUT ids.
int RefEncodeDcSection(const pikgpu_frame* f, std::string* out) {
  AcStrategyImage acs;
  if (ImportAcStrategy(f, &acs) != 0) return -1;
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  Image3S dc(bx, by);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < by; ++y) {
      memcpy(dc.PlaneRow(c, y), f->dc[c] + y * f->dc_stride, bx * 2);
    }
  }
  ImageI qf(bx, by);
  ImageB luts(bx, by);
  for (size_t y = 0; y < by; ++y) {
    memcpy(qf.Row(y), f->quant_field + y * f->quant_field_stride, bx * 4);
    memcpy(luts.Row(y), f->sigma_lut_ids + y * f->sigma_lut_stride, bx);
  }
  const size_t gxs = DivCeil<size_t>(bx, kDcGroupDimInBlocks);
  const size_t gys = DivCeil<size_t>(by, kDcGroupDimInBlocks);
  std::vector<std::string> codes;
  for (size_t gy = 0; gy < gys; ++gy) {
    for (size_t gx = 0; gx < gxs; ++gx) {
      const Rect rect(gx * kDcGroupDimInBlocks, gy * kDcGroupDimInBlocks,
                      kDcGroupDimInBlocks, kDcGroupDimInBlocks, bx, by);
      Image3S residuals(rect.xsize(), rect.ysize());
      ShrinkDC(rect, dc, &residuals);
      const std::string dc_code =
          EncodeImageData(Rect(residuals), residuals, nullptr);
      std::vector<std::vector<Token>> tokens(1);
      TokenizeAcStrategy(rect, acs, nullptr, &tokens[0]);
      TokenizeQuantField(rect, qf, nullptr, acs, &tokens[0]);
      TokenizeARParameters(rect, luts, acs, &tokens[0]);
      std::vector<uint8_t> context_map;
      std::vector<ANSEncodingData> codes_ans;
      const std::string histo = BuildAndEncodeHistograms(
          kNumControlFieldContexts, tokens, &codes_ans, &context_map, nullptr);
      codes.push_back(dc_code + histo +
                      WriteTokens(tokens[0], codes_ans, context_map, nullptr));
    }
  }
  std::vector<uint8_t> toc(
      DivCeil<size_t>(U32Coder::MaxEncodedBits(kDcGroupSizeDistribution) *
                          codes.size(), 8) + 8, 0);
  size_t pos = 0;
  for (const std::string& c : codes) {
    if (!U32Coder::Write(kDcGroupSizeDistribution, c.size(), &pos,
                         toc.data())) {
      return Fail("DC group too large for the size table");
    }
  }
  WriteZeroesToByteBoundary(&pos, toc.data());
  out->assign(reinterpret_cast<const char*>(toc.data()), pos / 8);
  for (const std::string& c : codes) *out += c;
  return 0;
}

// DecodeDCGroups' reading up to DequantDC from `reader` (positioned at the
// size table, as the frame decoder leaves it) over `data` (the whole
// buffer, which the group readers index absolutely); the reader ends past
// the groups.
int RefDecodeDcSection(BitReader* reader, const uint8_t* data, size_t size,
                       size_t bx, size_t by, AcStrategyImage* acs, ImageI* qf,
                       ImageB* luts, Image3S* dc) {
  const size_t gxs = DivCeil<size_t>(bx, kDcGroupDimInBlocks);
  const size_t gys = DivCeil<size_t>(by, kDcGroupDimInBlocks);
  const size_t ng = gxs * gys;
  std::vector<size_t> offsets(1, 0);
  for (size_t i = 0; i < ng; ++i) {
    offsets.push_back(offsets.back() +
                      U32Coder::Read(kDcGroupSizeDistribution, reader));
  }
  if (!reader->JumpToByteBoundary()) return Fail("size table padding");
  const size_t begin = reader->Position();
  reader->SkipBits(offsets.back() * kBitsPerByte);
  if (reader->Position() > size) return Fail("groups past the section");
  *acs = AcStrategyImage(bx, by);
  *qf = ImageI(bx, by);
  ZeroFillImage(qf);
  *luts = ImageB(bx, by);
  ZeroFillImage(luts);
  *dc = Image3S(bx, by);
  GroupDecCache cache;
  for (size_t gi = 0; gi < ng; ++gi) {
    const size_t gx = gi % gxs, gy = gi / gxs;
    const Rect rect(gx * kDcGroupDimInBlocks, gy * kDcGroupDimInBlocks,
                    kDcGroupDimInBlocks, kDcGroupDimInBlocks, bx, by);
    BitReader br(data, begin + offsets[gi + 1]);
    br.SkipBits((begin + offsets[gi]) * kBitsPerByte);
    cache.InitDecodeDCGroup(rect.xsize(), rect.ysize());
    if (!DecodeImage(&br, Rect(cache.quantized_dc), &cache.quantized_dc)) {
      return Fail("DecodeImage failed");
    }
    ExpandDC(Rect(cache.quantized_dc), &cache.quantized_dc, &cache.dc_y,
             &cache.dc_xz_residuals, &cache.dc_xz_expanded);
    if (!br.JumpToByteBoundary()) return Fail("DC padding");
    ANSCode code;
    std::vector<uint8_t> context_map;
    if (!DecodeHistograms(&br, kNumControlFieldContexts, 256, &code,
                          &context_map)) {
      return Fail("control field histograms");
    }
    if (!br.JumpToByteBoundary()) return Fail("histogram padding");
    ANSSymbolReader decoder(&code);
    if (!DecodeAcStrategy(&br, &decoder, context_map, &cache.ac_strategy_raw,
                          rect, acs, nullptr)) {
      return Fail("DecodeAcStrategy failed");
    }
    if (!DecodeQuantField(&br, &decoder, context_map, rect, *acs, qf,
                          nullptr)) {
      return Fail("DecodeQuantField failed");
    }
    if (!DecodeARParameters(&br, &decoder, context_map, rect, *acs, luts)) {
      return Fail("DecodeARParameters failed");
    }
    if (!decoder.CheckANSFinalState()) return Fail("ANS final state");
    if (!br.JumpToByteBoundary()) return Fail("control field padding");
    for (int c = 0; c < 3; ++c) {
      for (size_t y = 0; y < rect.ysize(); ++y) {
        memcpy(dc->PlaneRow(c, rect.y0() + y) + rect.x0(),
               cache.quantized_dc.ConstPlaneRow(c, y),
               rect.xsize() * sizeof(int16_t));
      }
    }
  }
  return 0;
}

}  // namespace

// The section (size table + groups) of the frame's quantized DC, strategy,
// quant field and sigma LUT ids into `out` (may be NULL to size it).
API int pikref_encode_dc_groups(const pikgpu_frame* f, uint8_t* out,
                                size_t cap, size_t* total) {
  if (f == nullptr) return Fail("null frame");
  std::string sec;
  if (RefEncodeDcSection(f, &sec) != 0) return -1;
  if (total != nullptr) *total = sec.size();
  if (out == nullptr) return 0;
  if (sec.size() > cap) return Fail("output buffer too small");
  memcpy(out, sec.data(), sec.size());
  return 0;
}

// Decodes a DC groups section for an xsize x ysize frame into the
// quantized DC (after ExpandDC), the combined strategy bytes
// (AcStrategyImage's layers), the quant field and the sigma LUT ids;
// *consumed receives the section's length.
API int pikref_decode_dc_groups(uint32_t xsize, uint32_t ysize,
                                const uint8_t* data, size_t size,
                                int16_t* const* dc_out, uint8_t* acs_out,
                                int32_t* qf_out, uint8_t* lut_out,
                                size_t* consumed) {
  const size_t bx = DivCeil<size_t>(xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(ysize, kBlockDim);
  PaddedBytes compressed(size);
  if (size != 0) memcpy(compressed.data(), data, size);
  BitReader reader(compressed.data(), compressed.size());
  AcStrategyImage acs;
  ImageI qf;
  ImageB luts;
  Image3S dc;
  if (RefDecodeDcSection(&reader, compressed.data(), size, bx, by, &acs, &qf,
                         &luts, &dc) != 0) {
    return -1;
  }
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < by; ++y) {
      memcpy(dc_out[c] + y * bx, dc.ConstPlaneRow(c, y), bx * 2);
    }
  }
  for (size_t y = 0; y < by; ++y) {
    memcpy(acs_out + y * bx, acs.ConstRaw().ConstRow(y), bx);
    memcpy(qf_out + y * bx, qf.ConstRow(y), bx * 4);
    memcpy(lut_out + y * bx, luts.ConstRow(y), bx);
  }
  if (consumed != nullptr) *consumed = reader.Position();
  return 0;
}

// --------------------------------------------------------------------------
// Whole frames: PixelsToPikPass's bitstream assembly (pik/pik_frame.cc:
// 555-681) over the given frame with the reference's own encoders, and the
// decoding half of PikPassToPixels / PikGroupToPixels (pik/pik_frame.cc:
// 846-931, 935-1110) with its own decoders, restated over public pieces
// because the drivers link colour management and the lossless codecs.

namespace {

constexpr uint32_t kGroupSizeDistribution = 0x150F0E0C;

// Bits of pikref_encode_frame's `variant`.
enum {
  kVarFrameInfo = 1,      // a non-default FrameInfo (duration, timecode)
  kVarGroupHeaders = 2,   // odd groups: a tile header with a projective
                          // transform, group extensions
  kVarStaticDots = 4,     // a placement of the static patch
  kVarFastAc = 8,         // BuildAndEncodeHistogramsFast for the AC
  kVarDownsampling = 16,  // two downsampling_factor_to_passes entries
  kVarCopyTable = 32,     // eight encoded tables (the last a copy): two
                          // full table sets
  kVarExtensions = 64     // FrameHeader extensions mask set
};

std::string BitsToString(const std::vector<uint8_t>& buf, size_t bits) {
  return std::string(reinterpret_cast<const char*>(buf.data()), bits / 8);
}

// GroupHeader (pik/headers.h:106-136) for frames without alpha, through
// the reference's own field visitors: the real bundle's Alpha member
// instantiates the Brotli byte coder, which is not in the checkout. With
// nonserialized_have_alpha false the real bundle visits exactly these
// fields.
struct RefGroupHeader {
  RefGroupHeader() { Bundle::Init(this); }
  static const char* Name() { return "GroupHeader"; }
  template <class Visitor>
  Status VisitFields(Visitor* PIK_RESTRICT visitor) {
    if (visitor->AllDefault(*this, &all_default)) return true;
    for (size_t i = 0; i < kNumTilesPerGroup; ++i) {
      PIK_RETURN_IF_ERROR(visitor->VisitNested(&tile_headers[i]));
    }
    visitor->BeginExtensions(&extensions);
    return visitor->EndExtensions();
  }
  bool all_default;
  TileHeader tile_headers[kNumTilesPerGroup];
  uint64_t extensions;
};

}  // namespace

// Writes one frame: FrameHeader, quantization tables, quantizer, colour
// maps, DC groups, dot dictionary, dequant control field and map, then per
// pass the group size table and the groups (GroupHeader in the first
// pass). Pass 0's AC is f->ac; pass i > 0 takes extra_ac[3 (i - 1) + c]
// (rows of f->ac_stride). noise (may be NULL): num_passes x groups x
// {alpha, gamma, beta}.
API int pikref_encode_frame(const pikgpu_frame* f, int num_passes,
                            const int16_t* const* extra_ac, const float* noise,
                            int variant, uint8_t* out, size_t cap,
                            size_t* total) {
  if (f == nullptr || num_passes < 1) return Fail("bad arguments");
  AcStrategyImage acs;
  if (ImportAcStrategy(f, &acs) != 0) return -1;
  const size_t bx = DivCeil<size_t>(f->xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(f->ysize, kBlockDim);
  const size_t tx = DivCeil<size_t>(f->xsize, kColorTileDim);
  const size_t ty = DivCeil<size_t>(f->ysize, kColorTileDim);
  const size_t gxs = DivCeil<size_t>(f->xsize, kGroupDim);
  const size_t gys = DivCeil<size_t>(f->ysize, kGroupDim);
  const size_t ng = gxs * gys;

  FrameHeader fh;
  if (variant & kVarFrameInfo) {
    fh.frame.duration = 7;
    fh.frame.have_timecode = true;
    fh.frame.timecode = 0x01020304;
    fh.frame.is_keyframe = true;
  }
  fh.encoding = ImageEncoding::kPasses;
  fh.flags = f->flags;
  fh.gaborish = static_cast<GaborishStrength>(f->gaborish);
  fh.predict_lf = f->predict_lf != 0;
  fh.predict_hf = f->predict_hf != 0;
  fh.have_adaptive_reconstruction = f->adaptive_reconstruction != 0;
  fh.epf_params.enable_adaptive = f->epf_enable_adaptive != 0;
  fh.epf_params.use_sharpened = f->epf_use_sharpened != 0;
  fh.num_passes = num_passes;
  if (variant & kVarDownsampling) {
    fh.downsampling_factor_to_passes = {{2, 0}, {8, 5}};
  }
  if (variant & kVarExtensions) fh.extensions = 1;
  size_t ext_bits = 0, total_bits = 0;
  if (!CanEncode(fh, &ext_bits, &total_bits)) return Fail("frame header");
  std::vector<uint8_t> hdr(DivCeil<size_t>(total_bits, 8) + 8, 0);
  size_t pos = 0;
  if (!WritePassHeader(fh, ext_bits, &pos, hdr.data())) {
    return Fail("WritePassHeader failed");
  }
  WriteZeroesToByteBoundary(&pos, hdr.data());
  std::string bytes = BitsToString(hdr, pos);

  DequantMatrices matrices(/*need_inv_matrices=*/false);
  if (variant & kVarCopyTable) {
    matrices.encodings_.assign(8, QuantEncoding::Library(0));
    matrices.encodings_[7] = QuantEncoding::Copy(3);
  }
  bytes += matrices.Encode(nullptr);
  const Quantizer quantizer = FrameQuantizer(f, &matrices);
  bytes += quantizer.Encode(nullptr);
  ImageI ytob(tx, ty), ytox(tx, ty);
  for (size_t y = 0; y < ty; ++y) {
    memcpy(ytob.Row(y), f->ytob_map + y * f->ytob_stride, tx * 4);
    memcpy(ytox.Row(y), f->ytox_map + y * f->ytox_stride, tx * 4);
  }
  bytes += EncodeColorMap(ytob, Rect(ytob), f->ytob_dc, nullptr);
  bytes += EncodeColorMap(ytox, Rect(ytox), f->ytox_dc, nullptr);
  std::string dc_section;
  if (RefEncodeDcSection(f, &dc_section) != 0) return -1;
  bytes += dc_section;

  BlockDictionary dots = ImportDots(*f);
  if (variant & kVarStaticDots) {
    // BlockDictionary::Encode restarts its scan of the sorted placements
    // at the first one for the transformed pass (pik/block_dictionary.cc:
    // 148-156), so it writes transformed placements only when there are
    // no plain ones: the static patch is placed plain, or transformed in a
    // frame with only transformed dots.
    const size_t id = dots.dictionary_.size();
    bool plain = false;
    for (const BlockPosition& p : dots.positions_) plain |= !p.transform;
    std::vector<BlockPosition> positions = dots.positions_;
    if (plain || positions.empty()) {
      positions.push_back(BlockPosition(8, 16, id));
    } else {
      BlockPosition p(20, 24, id, 5, 3, 4);
      p.width = 4;
      positions.push_back(p);
    }
    dots = BlockDictionary(dots.dictionary_, positions);
  }
  bytes += dots.Encode(nullptr);
  ImageB dcf(tx, ty);
  for (size_t y = 0; y < ty; ++y) {
    memcpy(dcf.Row(y), f->dequant_cf + y * f->dequant_cf_stride, tx);
  }
  uint8_t table_map[kMaxQuantControlFieldValue][256];
  memcpy(table_map, f->dequant_map, sizeof(table_map));
  bytes += EncodeDequantControlField(dcf, nullptr);
  bytes += EncodeDequantControlFieldMap(quantizer.RawQuantField(), dcf,
                                        table_map, nullptr);

  for (int i = 0; i < num_passes; ++i) {
    const int16_t* const* planes = i == 0 ? f->ac : extra_ac + 3 * (i - 1);
    std::vector<std::string> codes(ng);
    for (size_t gi = 0; gi < ng; ++gi) {
      const size_t gx = gi % gxs, gy = gi / gxs;
      const Rect group(gx * kGroupDim, gy * kGroupDim, kGroupDim, kGroupDim,
                       bx * kBlockDim, by * kBlockDim);
      const Rect blocks(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                        kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
      if (i == 0) {
        RefGroupHeader gh;
        if ((variant & kVarGroupHeaders) && gi % 2 == 1) {
          gh.tile_headers[3].have_projective_transform = true;
          for (size_t k = 0; k < kNumProjectiveTransformParams; ++k) {
            gh.tile_headers[3].projective_transform_params.corner_coords[k] =
                uint32_t(k + 1);
          }
          // (the reference's field writer supports extensions in the
          // outermost bundle only)
          gh.extensions = 2;
        }
        size_t eb = 0, tb = 0;
        if (!Bundle::CanEncode(gh, &eb, &tb)) return Fail("group header");
        std::vector<uint8_t> buf(DivCeil<size_t>(tb, 8) + 8, 0);
        size_t gpos = 0;
        if (!Bundle::Write(gh, eb, &gpos, buf.data())) {
          return Fail("WriteGroupHeader failed");
        }
        WriteZeroesToByteBoundary(&gpos, buf.data());
        codes[gi] = BitsToString(buf, gpos);
      }
      EncCache cache;
      cache.xsize_blocks = blocks.xsize();
      cache.ysize_blocks = blocks.ysize();
      cache.ac_strategy = acs.Copy(blocks);
      cache.ac = Image3S(blocks.xsize() * 64, blocks.ysize());
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < blocks.ysize(); ++y) {
          memcpy(cache.ac.PlaneRow(c, y),
                 planes[c] + (blocks.y0() + y) * f->ac_stride +
                     blocks.x0() * 64,
                 blocks.xsize() * 64 * 2);
        }
      }
      NoiseParams np;
      if (noise != nullptr) {
        const float* n = noise + (size_t(i) * ng + gi) * 3;
        np.alpha = n[0];
        np.gamma = n[1];
        np.beta = n[2];
      }
      const PaddedBytes ac = EncodeToBitstream(
          cache, Rect(0, 0, group.xsize(), group.ysize()),
          quantizer.Copy(blocks), np, (variant & kVarFastAc) != 0, nullptr);
      codes[gi].append(reinterpret_cast<const char*>(ac.data()), ac.size());
    }
    std::vector<uint8_t> toc(
        DivCeil<size_t>(U32Coder::MaxEncodedBits(kGroupSizeDistribution) * ng,
                        8) + 8, 0);
    size_t tpos = 0;
    for (const std::string& c : codes) {
      if (!U32Coder::Write(kGroupSizeDistribution, c.size(), &tpos,
                           toc.data())) {
        return Fail("group too large for the size table");
      }
    }
    WriteZeroesToByteBoundary(&tpos, toc.data());
    bytes += BitsToString(toc, tpos);
    for (const std::string& c : codes) bytes += c;
  }
  if (total != nullptr) *total = bytes.size();
  if (out == nullptr) return 0;
  if (bytes.size() > cap) return Fail("output buffer too small");
  memcpy(out, bytes.data(), bytes.size());
  return 0;
}

// Decodes one frame with the reference's decoders in PikPassToPixels'
// order. hdr[16] receives flags, gaborish, predict_lf, predict_hf,
// adaptive reconstruction, epf enable_adaptive, epf sigma, use_sharpened,
// num_passes, global_scale, quant_dc, ytox_dc, ytob_dc, the number of full
// table sets, the number of dots and FrameHeader::size (low 31 bits). The
// per-block and per-tile arrays are row-major without padding; ac holds
// up to max_passes passes of 3 planes; dots (up to dots_cap) are the
// placements in decoding order with their patches (placements of an
// empty patch without transform, which change nothing, are left out).
// Where the reference would index outside its own tables (a dequant
// control field value >= 16, a quant value outside [1, 256], a dequant map
// entry naming a missing table, a dot position that wrapped below zero, a
// degenerate dot transform, no passes) this fails instead.
API int pikref_decode_frame(const uint8_t* data, size_t size, uint32_t xsize,
                            uint32_t ysize, int32_t* hdr, int16_t* const* dc,
                            uint8_t* acs_out, int32_t* qf_out,
                            uint8_t* lut_out, int32_t* ytox_out,
                            int32_t* ytob_out, uint8_t* dcf_out,
                            uint8_t* dqmap_out, pikgpu_dot* dots,
                            size_t dots_cap, int16_t* ac, size_t max_passes,
                            float* noise, size_t* consumed) {
  const size_t bx = DivCeil<size_t>(xsize, kBlockDim);
  const size_t by = DivCeil<size_t>(ysize, kBlockDim);
  const size_t tx = DivCeil<size_t>(xsize, kColorTileDim);
  const size_t ty = DivCeil<size_t>(ysize, kColorTileDim);
  const size_t gxs = DivCeil<size_t>(xsize, kGroupDim);
  const size_t gys = DivCeil<size_t>(ysize, kGroupDim);
  const size_t ng = gxs * gys, nb = bx * by;
  PaddedBytes compressed(size);
  if (size != 0) memcpy(compressed.data(), data, size);
  BitReader reader(compressed.data(), size);
  FrameHeader fh;
  if (!ReadPassHeader(&reader, &fh)) return Fail("ReadPassHeader failed");
  if (!reader.JumpToByteBoundary()) return Fail("frame header padding");
  if (fh.encoding != ImageEncoding::kPasses) return Fail("not kPasses");
  if (fh.has_alpha) return Fail("alpha");
  if (fh.num_passes == 0) return Fail("no passes");
  if (fh.num_passes > max_passes) return Fail("more passes than buffers");
  if (fh.flags & FrameHeader::kGradientMap) return Fail("gradient map");
  DequantMatrices matrices(/*need_inv_matrices=*/false);
  if (!matrices.Decode(&reader)) return Fail("DequantMatrices::Decode");
  Quantizer quantizer(&matrices, 0, 0);
  quantizer.Decode(&reader);
  if (!reader.JumpToByteBoundary()) return Fail("quantizer padding");
  ColorCorrelationMap cmap(xsize, ysize);
  DecodeColorMap(&reader, &cmap.ytob_map, &cmap.ytob_dc);
  DecodeColorMap(&reader, &cmap.ytox_map, &cmap.ytox_dc);
  AcStrategyImage acs;
  ImageI qf;
  ImageB luts;
  Image3S dcq;
  if (RefDecodeDcSection(&reader, compressed.data(), size, bx, by, &acs, &qf,
                         &luts, &dcq) != 0) {
    return -1;
  }
  BlockDictionary dict;
  if (!dict.Decode(&reader, bx * kBlockDim, by * kBlockDim)) {
    return Fail("BlockDictionary::Decode failed");
  }
  ImageB dcf(tx, ty);
  if (!DecodeDequantControlField(&reader, &dcf)) {
    return Fail("DecodeDequantControlField failed");
  }
  for (size_t y = 0; y < ty; ++y) {
    for (size_t x = 0; x < tx; ++x) {
      if (dcf.Row(y)[x] >= kMaxQuantControlFieldValue) {
        return Fail("dequant control field value >= 16");
      }
    }
  }
  for (size_t y = 0; y < by; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      const int q = qf.Row(y)[x];
      if (q < 1 || q > 256) return Fail("quant value outside [1, 256]");
    }
  }
  uint8_t table_map[kMaxQuantControlFieldValue][256];
  if (!DecodeDequantControlFieldMap(&reader, qf, dcf, table_map)) {
    return Fail("DecodeDequantControlFieldMap failed");
  }
  const size_t num_sets = matrices.encodings_.size() / kNumQuantKinds;
  for (size_t i = 0; i < kMaxQuantControlFieldValue; ++i) {
    for (size_t j = 0; j < 256; ++j) {
      if (table_map[i][j] >= num_sets) return Fail("missing table");
    }
  }
  size_t num_dots = 0;
  for (const BlockPosition& p : dict.positions_) {
    const QuantizedBlock& b =
        p.id < dict.dictionary_.size()
            ? dict.dictionary_[p.id]
            : kStaticBlocks[p.id - dict.dictionary_.size()];
    if (p.x >= bx * kBlockDim || p.y >= by * kBlockDim) {
      return Fail("dot position wrapped");
    }
    if (p.transform && (b.xsize == 0 || b.ysize == 0 ||
                        (p.dx == 0 && p.dy == 0) || p.width <= 0)) {
      return Fail("degenerate dot transform");
    }
    if (!p.transform && (b.xsize == 0 || b.ysize == 0)) continue;
    if (num_dots < dots_cap) {
      pikgpu_dot& d = dots[num_dots];
      memset(&d, 0, sizeof(d));
      d.x = uint32_t(p.x);
      d.y = uint32_t(p.y);
      d.xsize = uint32_t(b.xsize);
      d.ysize = uint32_t(b.ysize);
      d.transform = p.transform;
      d.dx = int32_t(p.dx);
      d.dy = int32_t(p.dy);
      d.width = int32_t(p.width);
      for (int c = 0; c < 3; ++c) {
        memcpy(d.pixels[c], b.pixels[c], b.xsize * b.ysize);
      }
    }
    ++num_dots;
  }

  std::vector<size_t> begin(fh.num_passes);
  std::vector<std::vector<size_t>> offsets(fh.num_passes);
  for (size_t i = 0; i < fh.num_passes; ++i) {
    offsets[i].push_back(0);
    for (size_t g = 0; g < ng; ++g) {
      offsets[i].push_back(offsets[i].back() +
                           U32Coder::Read(kGroupSizeDistribution, &reader));
    }
    if (!reader.JumpToByteBoundary()) return Fail("group size padding");
    begin[i] = reader.Position();
    reader.SkipBits(offsets[i].back() * kBitsPerByte);
    if (reader.Position() > size) return Fail("groups past the end");
  }
  for (size_t g = 0; g < ng; ++g) {
    const size_t gx = g % gxs, gy = g / gxs;
    const Rect blocks(gx * kGroupDimInBlocks, gy * kGroupDimInBlocks,
                      kGroupDimInBlocks, kGroupDimInBlocks, bx, by);
    for (size_t i = 0; i < fh.num_passes; ++i) {
      BitReader br(compressed.data(), begin[i] + offsets[i][g + 1]);
      br.SkipBits((begin[i] + offsets[i][g]) * kBitsPerByte);
      if (i == 0) {
        RefGroupHeader gh;
        if (!Bundle::Read(&br, &gh)) return Fail("ReadGroupHeader failed");
        if (!br.JumpToByteBoundary()) return Fail("group header padding");
      }
      Image3S group_ac(blocks.xsize() * kDCTBlockSize, blocks.ysize());
      NoiseParams np;
      if (RefDecodeAcGroup(&br, acs, blocks, &group_ac, &np) != 0) return -1;
      if (!br.JumpToByteBoundary()) return Fail("pass padding");
      int16_t* pass_ac = ac + i * 3 * nb * 64;
      for (int c = 0; c < 3; ++c) {
        for (size_t y = 0; y < blocks.ysize(); ++y) {
          memcpy(pass_ac + c * nb * 64 + (blocks.y0() + y) * bx * 64 +
                     blocks.x0() * 64,
                 group_ac.ConstPlaneRow(c, y),
                 blocks.xsize() * 64 * sizeof(int16_t));
        }
      }
      float* n = noise + (i * ng + g) * 3;
      n[0] = np.alpha;
      n[1] = np.gamma;
      n[2] = np.beta;
    }
  }

  hdr[0] = int32_t(fh.flags);
  hdr[1] = int32_t(fh.gaborish);
  hdr[2] = fh.predict_lf;
  hdr[3] = fh.predict_hf;
  hdr[4] = fh.have_adaptive_reconstruction;
  hdr[5] = fh.epf_params.enable_adaptive;
  hdr[6] = int32_t(fh.epf_params.sigma);
  hdr[7] = fh.epf_params.use_sharpened;
  hdr[8] = int32_t(fh.num_passes);
  {
    // global_scale has no accessor: read it back from the encoding
    const std::string q = quantizer.Encode(nullptr);
    hdr[9] = ((uint8_t(q[0]) << 8) | uint8_t(q[1])) + 1;
  }
  hdr[10] = quantizer.QuantDC();
  hdr[11] = cmap.ytox_dc;
  hdr[12] = cmap.ytob_dc;
  hdr[13] = int32_t(num_sets);
  hdr[14] = int32_t(num_dots);
  hdr[15] = int32_t(fh.size & 0x7FFFFFFF);
  for (int c = 0; c < 3; ++c) {
    for (size_t y = 0; y < by; ++y) {
      memcpy(dc[c] + y * bx, dcq.ConstPlaneRow(c, y), bx * 2);
    }
  }
  for (size_t y = 0; y < by; ++y) {
    memcpy(acs_out + y * bx, acs.ConstRaw().ConstRow(y), bx);
    memcpy(qf_out + y * bx, qf.ConstRow(y), bx * 4);
    memcpy(lut_out + y * bx, luts.ConstRow(y), bx);
  }
  for (size_t y = 0; y < ty; ++y) {
    memcpy(ytox_out + y * tx, cmap.ytox_map.ConstRow(y), tx * 4);
    memcpy(ytob_out + y * tx, cmap.ytob_map.ConstRow(y), tx * 4);
    memcpy(dcf_out + y * tx, dcf.ConstRow(y), tx);
  }
  memcpy(dqmap_out, table_map, sizeof(table_map));
  if (consumed != nullptr) *consumed = reader.Position();
  return 0;
}
