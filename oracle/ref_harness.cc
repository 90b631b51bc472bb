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
#include "pik/color_correlation.h"
#include "pik/compressed_dc.h"
#include "pik/compressed_image.h"
#include "pik/data_parallel.h"
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
