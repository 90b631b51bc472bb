// C ABI of the gfx950 reconstruction (include/pikgpu.h): context, frame
// upload with validation, launch sequence, downloads and stage dumps.
//
// The launch sequence mirrors the reference decoder's order: DequantDC and
// AdaptiveDCReconstruction (DecodeDCGroups, pik/compressed_dc.cc:436-528),
// the per-group prediction set-up of ReconOpsinImage
// (pik/compressed_image.cc:936-1084), the per-block dequant/prediction/IDCT
// (pik/compressed_image.cc:1086-1097), then FinalizeFrameDecoding
// (pik/compressed_image.cc:1144-1181).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "host_api.h"
#include "kernels.h"

using namespace pikgpu;

namespace {

// Kernel families timed by pikgpu_run (events ev[0..kNumRunTimers]), by
// pikgpu_fwd_adj (adj_ev[0..2]) and by pikgpu_quantize (enc_ev[0..1]).
enum Timer {
  kTDc, kTLf, kTBlocks, kTMinMax, kTEpf, kTDots, kTFinal, kNumRunTimers,
  kTAdjForward = kNumRunTimers, kTAdjBackward, kTEncQuantize, kNumTimers
};
const char* kTimerNames[kNumTimers] = {"dc",     "lf_predict", "blocks_idct",
                                       "guide",  "epf_ar",     "dots",
                                       "final_opsin", "adj_forward",
                                       "adj_backward", "enc_quantize"};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};


// Most AC passes of a frame (pikgpu_frame_add_pass).
constexpr int kMaxPasses = 16;

}  // namespace

struct pikgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  DeviceConsts* consts = nullptr;
  HostTables host;
  // device buffers (grown on demand)
  DevBuf ac, dc16, qf, acs, lut, dcf, ytox, ytob, dqmap, table, offsets, dots,
      dotbins, ac_extra;
  DevBuf desc, dcq, dcv, llf, tmp2x2, lf2x2, pred2x2, coeffs, idct, guide, epf, ar,
      dots_img, gab, out, mm_partial, mm, ar_keep, srgb8, gimg, predlist;
  // pikgpu_fwd_adj staging (host-pointer calls) and loss partials
  DevBuf adj_qf, adj_in, adj_lin, adj_dqf, adj_dc, adj_loss;
  // pikgpu_quantize: XYB input, coefficients, DC, the inverse dequant table
  DevBuf eopsin, ecoef, edcf, itable, ecorner;
  // host storage of pikgpu_decode_frame's parse, reused frame to frame
  pikgpu_parsed* parsed = nullptr;
  // pikgpu_search_quant_field: log field, Adam moments, weights, per-block
  // distortion, rate partials, the rate weight, loss / rate histories
  DevBuf srch;
  bool enc_map_ok = false;  // every quant_table index of the frame in range
  bool has_lines = false;   // a LINES block (no adjoint: fwd_adj, search)
  int global_scale = 0, quant_dc = 0;  // the uploaded frame's
  float dc_matrix0[3] = {0, 0, 0};     // DequantMatrix(0, DCT8, c)[0]
  bool adj_ready = false;  // LLF / LF caches hold the uploaded frame's
  bool adj_timed = false;
  bool uploaded = false;
  bool ran = false;
  bool kept = false;
  int out_flags = kOutLinear;
  int downsample = 1;
  int ran_flags = 0;
  int spitch = 0;
  int out_xsize = 0, out_ysize = 0;  // output size of the last run
  bool ran_downsampled = false;
  bool timed = false;   // the last run recorded timing events
  FrameParams P;
  FrameBuffers B;
  int num_tables = 0;
  hipEvent_t ev[kNumRunTimers + 1] = {};
  hipEvent_t adj_ev[3] = {};
  hipEvent_t enc_ev[2] = {};
  float timings[kNumTimers] = {};
  // Pinned host staging for strided host planes with short rows (packed
  // here, then one linear copy): [0, stage_used) may still be in flight
  // until stage_ev, recorded after the last staged copy, has completed.
  void* stage = nullptr;
  size_t stage_cap = 0, stage_used = 0;
  hipEvent_t stage_ev = nullptr;
};

namespace {

int Fail(pikgpu_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                  \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      return Fail(ctx, PIKGPU_EHIP,                                         \
                  std::string(#expr) + ": " + hipGetErrorString(e_));       \
    }                                                                       \
  } while (0)

int Grow(pikgpu_ctx* ctx, DevBuf* b, size_t bytes) {
  if (bytes == 0) bytes = 4;
  if (b->cap >= bytes) return PIKGPU_OK;
  if (b->p) (void)hipFree(b->p);
  b->p = nullptr;
  b->cap = 0;
  hipError_t e = hipMalloc(&b->p, bytes);
  if (e != hipSuccess) {
    return Fail(ctx, PIKGPU_ENOMEM,
                std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  b->cap = bytes;
  return PIKGPU_OK;
}

void FreeAll(pikgpu_ctx* ctx) {
  DevBuf* all[] = {&ctx->ac,      &ctx->ac_extra, &ctx->dc16, &ctx->qf,
                   &ctx->acs,
                   &ctx->lut,     &ctx->dcf,    &ctx->ytox,    &ctx->ytob,
                   &ctx->dqmap,   &ctx->table,  &ctx->offsets, &ctx->dots,
                   &ctx->dotbins, &ctx->srgb8,
                   &ctx->desc,    &ctx->dcq,     &ctx->dcv,    &ctx->llf,     &ctx->tmp2x2,
                   &ctx->lf2x2,   &ctx->pred2x2, &ctx->coeffs, &ctx->idct,
                   &ctx->guide,   &ctx->epf,    &ctx->ar,      &ctx->dots_img,
                   &ctx->gab,     &ctx->out,    &ctx->mm_partial, &ctx->mm,
                   &ctx->ar_keep, &ctx->gimg,   &ctx->predlist, &ctx->adj_qf, &ctx->adj_in, &ctx->adj_lin,
                   &ctx->adj_dqf, &ctx->adj_dc, &ctx->adj_loss,
                   &ctx->eopsin, &ctx->ecoef, &ctx->edcf, &ctx->itable,
                   &ctx->ecorner,
                   &ctx->srch};
  for (DevBuf* b : all) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->cap = 0;
  }
}

// Rows shorter than this are packed into the pinned staging buffer when
// the host plane is strided; longer ones go row by row (per-row copy cost
// is then small against the row's own transfer).
constexpr size_t kStageRowMax = 32 << 10;
// Most bytes staged at once (larger planes go in chunks of rows).
constexpr size_t kStageChunk = 16 << 20;

// Makes the pinned staging buffer at least n <= kStageChunk bytes with
// nothing in flight; it grows by doubling up to kStageChunk, so a context
// of small frames keeps a small one.
int StageReserve(pikgpu_ctx* ctx, size_t n) {
  if (ctx->stage_used > 0) {
    HIP_TRY(ctx, hipEventSynchronize(ctx->stage_ev));
    ctx->stage_used = 0;
  }
  if (n > ctx->stage_cap) {
    const size_t cap = std::min(kStageChunk, std::max(n, 2 * ctx->stage_cap));
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_cap = 0;
    HIP_TRY(ctx, hipHostMalloc(&ctx->stage, cap, 0));
    ctx->stage_cap = cap;
  }
  return PIKGPU_OK;
}

// Packs a strided host plane (rows shorter than kStageRowMax) into the
// context's pinned staging buffer and queues one linear copy per chunk of
// at most kStageChunk bytes.
int StageUpload(pikgpu_ctx* ctx, void* dst, const void* src,
                size_t src_stride_b, size_t row_b, size_t rows) {
  const size_t per = std::max<size_t>(1, kStageChunk / row_b);
  for (size_t y0 = 0; y0 < rows; y0 += per) {
    const size_t ny = std::min(per, rows - y0), n = row_b * ny;
    if (ctx->stage_used + n > ctx->stage_cap) {
      const int rc = StageReserve(ctx, n);
      if (rc != PIKGPU_OK) return rc;
    }
    uint8_t* p = static_cast<uint8_t*>(ctx->stage) + ctx->stage_used;
    const uint8_t* q = static_cast<const uint8_t*>(src) + y0 * src_stride_b;
    for (size_t y = 0; y < ny; ++y) {
      memcpy(p + y * row_b, q + y * src_stride_b, row_b);
    }
    HIP_TRY(ctx, hipMemcpyAsync(static_cast<uint8_t*>(dst) + y0 * row_b, p, n,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->stage_ev, ctx->stream));
    ctx->stage_used += (n + 255) & ~size_t(255);
  }
  return PIKGPU_OK;
}

// Copies a strided host plane into a dense device region: one linear copy
// when the plane is dense, packed through pinned staging when its rows are
// short, else the 2-D copy.
int Upload2D(pikgpu_ctx* ctx, void* dst, const void* src, size_t src_stride_b,
             size_t row_b, size_t rows) {
  if (rows > 1 && src_stride_b != row_b && row_b < kStageRowMax) {
    return StageUpload(ctx, dst, src, src_stride_b, row_b, rows);
  }
  HIP_TRY(ctx, Copy2D(dst, row_b, src, src_stride_b, row_b, rows,
                      hipMemcpyHostToDevice, ctx->stream));
  return PIKGPU_OK;
}

float ToF(double v) { return static_cast<float>(v); }

}  // namespace

namespace pikgpu {

hipError_t Copy2D(void* dst, size_t dpitch, const void* src, size_t spitch,
                  size_t row_b, size_t rows, hipMemcpyKind kind,
                  hipStream_t stream) {
  if (rows <= 1 || (dpitch == row_b && spitch == row_b)) {
    return hipMemcpyAsync(dst, src, row_b * rows, kind, stream);
  }
  return hipMemcpy2DAsync(dst, dpitch, src, spitch, row_b, rows, kind,
                          stream);
}

// Inverse of the opsin absorbance matrix as Inv3x3Matrix<float> computes it
// (pik/linalg.h:153-170, pik/opsin_params.cc:13-26).
void OpsinInverse(float out[9]) {
  const float kScale = 255.0f;
  float m[9] = {
      ToF(0.29956550340058319 / kScale), ToF(0.63373087833825936 / kScale),
      ToF(0.077705617820981968 / kScale), ToF(0.22158691104574774 / kScale),
      ToF(0.68491388044116142 / kScale), ToF(0.096254234043612538 / kScale),
      ToF(0.20062661225219422 / kScale), ToF(0.070366199217588729 / kScale),
      ToF(0.5571760754215358 / kScale)};
  float t[9];
  t[0] = m[4] * m[8] - m[5] * m[7];
  t[1] = m[2] * m[7] - m[1] * m[8];
  t[2] = m[1] * m[5] - m[2] * m[4];
  t[3] = m[5] * m[6] - m[3] * m[8];
  t[4] = m[0] * m[8] - m[2] * m[6];
  t[5] = m[2] * m[3] - m[0] * m[5];
  t[6] = m[3] * m[7] - m[4] * m[6];
  t[7] = m[1] * m[6] - m[0] * m[7];
  t[8] = m[0] * m[4] - m[1] * m[3];
  const float idet = 1.0 / (m[0] * t[0] + m[1] * t[3] + m[2] * t[6]);
  for (int i = 0; i < 9; ++i) out[i] = t[i] * idet;
}

}  // namespace pikgpu

namespace {

size_t DivCeil(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace

extern "C" {

pikgpu_ctx* pikgpu_create(int device, const pikgpu_config* config) {
  if (config != nullptr && config->abi_version != PIKGPU_ABI_VERSION) {
    return nullptr;
  }
  pikgpu_ctx* ctx = new pikgpu_ctx;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete ctx;
    return nullptr;
  }
  for (int i = 0; i <= kNumRunTimers + 3; ++i) {
    hipEvent_t* e = i <= kNumRunTimers ? &ctx->ev[i]
                                       : &ctx->adj_ev[i - kNumRunTimers - 1];
    if (hipEventCreate(e) != hipSuccess) {
      delete ctx;
      return nullptr;
    }
  }
  for (hipEvent_t& e : ctx->enc_ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return nullptr;
    }
  }
  if (hipEventCreateWithFlags(&ctx->stage_ev, hipEventDisableTiming) !=
      hipSuccess) {
    delete ctx;
    return nullptr;
  }
  // Constant tables.
  DeviceConsts* h = new DeviceConsts;
  memset(h, 0, sizeof(*h));
  FillScaleTables(&h->tscale[0][0][0], &h->ascale[0][0][0]);
  ComputeHostTables(&ctx->host);
  memcpy(h->mul_table, ctx->host.mul_table, sizeof(h->mul_table));
  OpsinInverse(h->opsin_inv);
  AdjointNorms(h->adj_norm);
  ComputeGrayXyb(&h->gray);
  if (hipMalloc(&ctx->consts, sizeof(DeviceConsts)) != hipSuccess ||
      hipMemcpy(ctx->consts, h, sizeof(DeviceConsts), hipMemcpyHostToDevice) !=
          hipSuccess) {
    delete h;
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return nullptr;
  }
  delete h;
  return ctx;
}

void pikgpu_destroy(pikgpu_ctx* ctx) {
  if (ctx == nullptr) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  FreeAll(ctx);
  pikgpu_parsed_free(ctx->parsed);
  if (ctx->consts) (void)hipFree(ctx->consts);
  for (int i = 0; i <= kNumRunTimers; ++i) {
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  }
  for (hipEvent_t e : ctx->adj_ev) {
    if (e) (void)hipEventDestroy(e);
  }
  for (hipEvent_t e : ctx->enc_ev) {
    if (e) (void)hipEventDestroy(e);
  }
  if (ctx->stage_ev) (void)hipEventDestroy(ctx->stage_ev);
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* pikgpu_last_error(const pikgpu_ctx* ctx) {
  if (ctx == nullptr) return "null context";
  return ctx->error.c_str();
}

void* pikgpu_stream(pikgpu_ctx* ctx) {
  return ctx ? static_cast<void*>(ctx->stream) : nullptr;
}

}  // extern "C"

namespace {

int VFail(std::string* err, int code, const char* msg) {
  if (err) *err = msg;
  return code;
}

}  // namespace

namespace pikgpu {

// Checks that a frame is complete and self-consistent (what the reference's
// decoder establishes while reading the bitstream; a violation there is a
// decoding error). Host-only; touches no device.
int ValidateFrame(const pikgpu_frame* f, std::string* err) {
  if (f == nullptr) return VFail(err, PIKGPU_EINVAL, "null frame");
  if (f->xsize == 0 || f->ysize == 0 || f->xsize > (1u << 16) ||
      f->ysize > (1u << 16)) {
    return VFail(err, PIKGPU_EINVAL, "invalid image size");
  }
  if (f->global_scale < 1 || f->global_scale > 32768 || f->quant_dc < 1 ||
      f->quant_dc > 256) {
    return VFail(err, PIKGPU_EINVAL, "invalid quantizer scalars");
  }
  if (f->gaborish < 0 || f->gaborish > 4) {
    return VFail(err, PIKGPU_EINVAL, "invalid gaborish strength");
  }
  if (f->flags & PIKGPU_FLAG_GRADIENT_MAP) {
    return VFail(err, PIKGPU_EUNSUPPORTED,
                 "gradient-map DC (FrameHeader::kGradientMap) is not supported");
  }
  if (f->flags & PIKGPU_FLAG_NOISE) {
    return VFail(err, PIKGPU_EUNSUPPORTED,
                 "noise synthesis (FrameHeader::kNoise) is not supported");
  }
  if (f->flags & ~uint32_t(PIKGPU_FLAG_GRADIENT_MAP | PIKGPU_FLAG_GRAYSCALE_OPT |
                           PIKGPU_FLAG_NOISE)) {
    return VFail(err, PIKGPU_EUNSUPPORTED, "unknown frame header flag");
  }
  if (f->adaptive_reconstruction && !f->epf_enable_adaptive) {
    return VFail(err, PIKGPU_EUNSUPPORTED,
                "fixed-sigma EPF (enable_adaptive = false) is not supported");
  }
  const size_t bx = DivCeil(f->xsize, 8), by = DivCeil(f->ysize, 8);
  const size_t tx = DivCeil(f->xsize, 64), ty = DivCeil(f->ysize, 64);
  if (!f->ac[0] || !f->ac[1] || !f->ac[2] || !f->dc[0] || !f->dc[1] ||
      !f->dc[2] || !f->quant_field || !f->ac_strategy || !f->sigma_lut_ids ||
      !f->dequant_cf || !f->ytox_map || !f->ytob_map || !f->dequant_map ||
      !f->dequant_table || !f->dequant_offsets) {
    return VFail(err, PIKGPU_EINVAL, "null array in frame");
  }
  if (f->ac_stride < bx * 64 || f->dc_stride < bx ||
      f->quant_field_stride < bx || f->ac_strategy_stride < bx ||
      f->sigma_lut_stride < bx || f->dequant_cf_stride < tx ||
      f->ytox_stride < tx || f->ytob_stride < tx) {
    return VFail(err, PIKGPU_EINVAL, "stride smaller than row");
  }
  if (f->num_dequant_offsets == 0 ||
      f->num_dequant_offsets % (kNumQuantKinds * 3) != 0) {
    return VFail(err, PIKGPU_EINVAL, "dequant offsets not a multiple of 21");
  }
  const int num_tables = int(f->num_dequant_offsets / (kNumQuantKinds * 3));
  // Every table entry a kind can address must lie in the table.
  const int kind_size[kNumQuantKinds] = {64, 64, 64, 64, 256, 1024, 64};
  for (int t = 0; t < num_tables; ++t) {
    for (int k = 0; k < kNumQuantKinds; ++k) {
      for (int c = 0; c < 3; ++c) {
        const size_t o = f->dequant_offsets[(t * kNumQuantKinds + k) * 3 + c];
        if (o + kind_size[k] > f->dequant_table_size) {
          return VFail(err, PIKGPU_EINVAL, "dequant offset out of table");
        }
      }
    }
  }
  for (int i = 0; i < 16 * 256; ++i) {
    if (f->dequant_map[i] >= num_tables) {
      return VFail(err, PIKGPU_EINVAL, "dequant_map refers to missing table");
    }
  }
  for (size_t y = 0; y < ty; ++y) {
    for (size_t x = 0; x < tx; ++x) {
      if (f->dequant_cf[y * f->dequant_cf_stride + x] >= 16) {
        return VFail(err, PIKGPU_EINVAL, "dequant control field >= 16");
      }
    }
  }
  // AC strategy: valid raw types, multi-block transforms complete, aligned
  // and inside the image (as AcStrategyImage::SetFromRaw establishes).
  for (size_t y = 0; y < by; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      const uint8_t a = f->ac_strategy[y * f->ac_strategy_stride + x];
      const int raw = a >> 4, idx = a & 15;
      const int qfv = f->quant_field[y * f->quant_field_stride + x];
      if (qfv < 1 || qfv > 256) {
        return VFail(err, PIKGPU_EINVAL, "quant field outside [1, 256]");
      }
      if (f->sigma_lut_ids[y * f->sigma_lut_stride + x] > 1) {
        return VFail(err, PIKGPU_EINVAL, "sigma lut id > 1");
      }
      if (raw > 8) return VFail(err, PIKGPU_EINVAL, "invalid AC strategy");
      const int s = AcsCovered(raw);
      if (idx >= s * s) return VFail(err, PIKGPU_EINVAL, "invalid block index");
      const size_t ox = x - idx % s, oy = y - idx / s;
      if (idx % s > int(x) || idx / s > int(y) || ox % s || oy % s ||
          ox + s > bx || oy + s > by) {
        return VFail(err, PIKGPU_EINVAL, "misaligned multi-block transform");
      }
      // A transform's first block checks the whole transform; every other
      // block only that its first block is one of the same kind, whose own
      // check then covers it.
      const int kmax = idx == 0 ? s * s : 1;
      for (int k = 0; k < kmax; ++k) {
        const uint8_t e =
            f->ac_strategy[(oy + k / s) * f->ac_strategy_stride + ox + k % s];
        if (e != ((raw << 4) | k)) {
          return VFail(err, PIKGPU_EINVAL, "inconsistent multi-block transform");
        }
      }
    }
  }
  for (size_t i = 0; i < f->num_dots; ++i) {
    const pikgpu_dot& d = f->dots[i];
    if (d.xsize < 1 || d.xsize > 8 || d.ysize < 1 || d.ysize > 8) {
      return VFail(err, PIKGPU_EINVAL, "dot size outside [1, 8]");
    }
    if (!d.transform && (d.x + d.xsize > bx * 8 || d.y + d.ysize > by * 8)) {
      return VFail(err, PIKGPU_EINVAL, "dot outside the image");
    }
    if (d.transform && (d.width <= 0 || (d.dx == 0 && d.dy == 0))) {
      return VFail(err, PIKGPU_EINVAL, "degenerate dot transform");
    }
  }

  return PIKGPU_OK;
}

// Geometry and derived scalars of a validated frame.
void FillParams(const pikgpu_frame* f, FrameParams* Pp) {
  FrameParams& P = *Pp;
  memset(&P, 0, sizeof(P));
  const size_t bx = DivCeil(f->xsize, 8), by = DivCeil(f->ysize, 8);
  const size_t tx = DivCeil(f->xsize, 64), ty = DivCeil(f->ysize, 64);
  P.xsize = f->xsize;
  P.ysize = f->ysize;
  P.bx = int(bx);
  P.by = int(by);
  P.px = int(bx * 8);
  P.py = int(by * 8);
  P.tx = int(tx);
  P.ty = int(ty);
  P.gx = int(DivCeil(bx, kGroupDimBlocks));
  P.gy = int(DivCeil(by, kGroupDimBlocks));
  P.predict_lf = f->predict_lf != 0;
  P.predict_hf = f->predict_hf != 0;
  P.gaborish = f->gaborish != 0;
  P.ar = f->adaptive_reconstruction != 0;
  P.use_sharpened = f->epf_use_sharpened != 0;
  P.num_dots = int(f->num_dots);
  P.downsample = 1;  // per run (pikgpu_set_downsample)
  P.num_passes = 1;  // more with pikgpu_frame_add_pass
  P.grayscale = f->grayscale != 0 || (f->flags & PIKGPU_FLAG_GRAYSCALE_OPT);
  // Quantizer::Decode (pik/quantizer.cc:63-71).
  P.inv_global_scale = static_cast<float>(1.0 * 65536 / f->global_scale);
  P.inv_quant_dc = P.inv_global_scale / f->quant_dc;
  for (int c = 0; c < 3; ++c) {
    const float m0 = f->dequant_table[f->dequant_offsets[c]];
    P.mul_dc[c] = m0 * P.inv_quant_dc;
    P.dc_half_step[c] = P.inv_quant_dc * m0 * 0.5f;
  }
  P.ytox_dc = float(f->ytox_dc - 128) * (1.0f / 256);
  P.ytob_dc = float(f->ytob_dc) * (1.0f / 128);
  P.dc_qf = f->quant_dc / 2;
  GaborishWeights(f->gaborish, P.gab_w);
  LfBlurWeights(P.lfblur_w);
  Ub4Kernel(P.ub4);
  const float kScale = 255.0f;
  P.neg_bias[0] = -ToF(0.26786006338144885 / kScale);
  P.neg_bias[1] = -ToF(0.24494032763907073 / kScale);
  P.neg_bias[2] = -ToF(0.14255999980363571 / kScale);
}

// First blocks grouped by the class of their pred2x2 computation
// (KPred2x2Interior): DCT8-like, 2x2 Hadamard, DCT16, DCT32; `end` receives
// the end of each class in `list`.
void PredClasses(const pikgpu_frame* f, std::vector<uint32_t>* list,
                 int end[4]) {
  const size_t bx = DivCeil(f->xsize, 8), by = DivCeil(f->ysize, 8);
  std::vector<uint32_t> plist[4];
  for (size_t y = 0; y < by; ++y) {
    const uint8_t* row = f->ac_strategy + y * f->ac_strategy_stride;
    for (size_t x = 0; x < bx; ++x) {
      if ((row[x] & 15) != 0) continue;
      const int raw = row[x] >> 4;
      const int cls = raw == kDCT16 ? 2
                      : raw == kDCT32 ? 3
                      : (raw == kDCT || raw == kDCTNoHF || raw == kLines) ? 0
                                                                           : 1;
      plist[cls].push_back(uint32_t(y * bx + x));
    }
  }
  list->clear();
  for (int k = 0; k < 4; ++k) {
    list->insert(list->end(), plist[k].begin(), plist[k].end());
    end[k] = int(list->size());
  }
}

// Bins the dots into 32x32 tiles of the padded image. Every dot goes to
// each tile its footprint may touch (transformed dots: the bounding box of
// the splat's four corners, widened by two pixels), and each tile's list
// keeps the order BlockDictionary::Apply uses (pik/block_dictionary.cc:
// 346-482): placements copied as-is first, then transformed ones, each in
// list order. Output: tile ids, CSR starts, dot indices.
void BinDots(const pikgpu_frame* f, int px, int py, std::vector<int32_t>* ids,
             std::vector<int32_t>* starts, std::vector<int32_t>* list) {
  const int tiles_x = int(DivCeil(px, 32)), tiles_y = int(DivCeil(py, 32));
  std::vector<std::pair<int64_t, int32_t>> entries;  // (tile<<32 | rank, dot)
  int64_t rank = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (size_t i = 0; i < f->num_dots; ++i) {
      const pikgpu_dot& d = f->dots[i];
      if ((d.transform != 0) != (pass == 1)) continue;
      double x0, x1, y0, y1;
      if (!d.transform) {
        x0 = d.x;
        y0 = d.y;
        x1 = double(d.x) + d.xsize;
        y1 = double(d.y) + d.ysize;
      } else {
        const double norm = std::sqrt(double(d.dx) * d.dx + double(d.dy) * d.dy);
        const double half = d.width * 0.5;
        const double ex = d.dy / norm * half, ey = -d.dx / norm * half;
        const double cx[4] = {0.0, double(d.dx), ex, d.dx + ex};
        const double cy[4] = {0.0, double(d.dy), ey, d.dy + ey};
        x0 = x1 = d.x;
        y0 = y1 = d.y;
        for (int k = 0; k < 4; ++k) {
          x0 = std::min(x0, d.x + cx[k]);
          x1 = std::max(x1, d.x + cx[k]);
          y0 = std::min(y0, d.y + cy[k]);
          y1 = std::max(y1, d.y + cy[k]);
        }
        x0 -= 2;
        y0 -= 2;
        x1 += 3;
        y1 += 3;
      }
      const int tx0 = int(std::max(0.0, std::floor(x0 / 32)));
      const int ty0 = int(std::max(0.0, std::floor(y0 / 32)));
      const int tx1 = int(std::min(double(tiles_x - 1), std::floor((x1 - 1) / 32)));
      const int ty1 = int(std::min(double(tiles_y - 1), std::floor((y1 - 1) / 32)));
      for (int ty = ty0; ty <= ty1; ++ty) {
        for (int tx = tx0; tx <= tx1; ++tx) {
          entries.emplace_back((int64_t(ty * tiles_x + tx) << 32) | rank,
                               int32_t(i));
        }
      }
      ++rank;
    }
  }
  std::sort(entries.begin(), entries.end());
  ids->clear();
  starts->clear();
  list->clear();
  for (size_t k = 0; k < entries.size(); ++k) {
    const int32_t tile = int32_t(entries[k].first >> 32);
    if (ids->empty() || ids->back() != tile) {
      ids->push_back(tile);
      starts->push_back(int32_t(k));
    }
    list->push_back(entries[k].second);
  }
  starts->push_back(int32_t(entries.size()));
}

}  // namespace pikgpu

extern "C" {

int pikgpu_default_dequant_tables(float* table, size_t table_cap,
                                  uint32_t* offsets, size_t offsets_cap,
                                  size_t* table_size, size_t* num_offsets) {
  const size_t n =
      DefaultDequantTables(table, table_cap, offsets, offsets_cap, num_offsets);
  if (table_size) *table_size = n;
  return PIKGPU_OK;
}

int pikgpu_validate_frame(const pikgpu_frame* frame, char* msg,
                          size_t msg_size) {
  std::string err;
  const int rc = ValidateFrame(frame, &err);
  if (msg && msg_size) {
    snprintf(msg, msg_size, "%s", rc == PIKGPU_OK ? "" : err.c_str());
  }
  return rc;
}

int pikgpu_frame_upload(pikgpu_ctx* ctx, const pikgpu_frame* f) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ctx->uploaded = false;
  ctx->ran = false;
  ctx->adj_ready = false;
  {
    std::string err;
    const int rc = ValidateFrame(f, &err);
    if (rc != PIKGPU_OK) return Fail(ctx, rc, err);
  }
  const size_t bx = DivCeil(f->xsize, 8), by = DivCeil(f->ysize, 8);
  const size_t tx = DivCeil(f->xsize, 64), ty = DivCeil(f->ysize, 64);
  const int num_tables = int(f->num_dequant_offsets / (kNumQuantKinds * 3));
  FrameParams& P = ctx->P;
  FillParams(f, &P);

  // --- buffers ---
  const size_t nblk = bx * by;
  const size_t ngroups = size_t(P.gx) * P.gy;
  const size_t ipitch = DivCeil(bx * 8, 32) * 32;
  const size_t opitch = DivCeil(f->xsize, 32) * 32;
  int rc;
#define GROW(buf, bytes) \
  if ((rc = Grow(ctx, &ctx->buf, (bytes))) != PIKGPU_OK) return rc
  GROW(ac, 3 * nblk * 64 * sizeof(int16_t));
  GROW(dc16, 3 * nblk * sizeof(int16_t));
  GROW(qf, nblk * 4);
  GROW(acs, nblk);
  GROW(lut, nblk);
  GROW(dcf, tx * ty);
  GROW(ytox, tx * ty * 4);
  GROW(ytob, tx * ty * 4);
  GROW(dqmap, 16 * 256);
  GROW(table, f->dequant_table_size * 4);
  GROW(offsets, f->num_dequant_offsets * 4);
  GROW(dots, f->num_dots * sizeof(pikgpu_dot));
  GROW(desc, 2 * nblk * 16);
  GROW(dcq, 3 * nblk * 4);
  GROW(dcv, 3 * nblk * 4);
  GROW(llf, ngroups * 3 * kLlfPlane * 4);
  GROW(tmp2x2, ngroups * 3 * kG2Plane * 4);
  GROW(lf2x2, ngroups * 3 * kG2Plane * 4);
  GROW(pred2x2, ngroups * 3 * kG2Plane * 4);
  GROW(idct, 3 * ipitch * by * 8 * 4);
  GROW(ar, 3 * ipitch * by * 8 * 4);
  // guide image: 6-pixel border, rows and pitch cover the 44-row, 44-byte
  // windows of every (possibly partial) 32x32 tile
  const size_t grows = DivCeil(by * 8, 32) * 32 + 16;
  const size_t gpitch = DivCeil(bx * 8, 32) * 32 + 64;
  GROW(gimg, 3 * grows * gpitch);
  GROW(out, 3 * opitch * f->ysize * 4);
  GROW(mm_partial, kMinMaxMaxGroups * 6 * 4);
  GROW(mm, 8 * 4);
#undef GROW

  // --- copies ---
  const size_t nb = nblk;
  for (int c = 0; c < 3; ++c) {
    if ((rc = Upload2D(ctx, static_cast<int16_t*>(ctx->ac.p) + c * nb * 64,
                       f->ac[c], f->ac_stride * 2, bx * 64 * 2, by)))
      return rc;
    if ((rc = Upload2D(ctx, static_cast<int16_t*>(ctx->dc16.p) + c * nb,
                       f->dc[c], f->dc_stride * 2, bx * 2, by)))
      return rc;
  }
  if ((rc = Upload2D(ctx, ctx->qf.p, f->quant_field, f->quant_field_stride * 4,
                     bx * 4, by)))
    return rc;
  if ((rc = Upload2D(ctx, ctx->acs.p, f->ac_strategy, f->ac_strategy_stride,
                     bx, by)))
    return rc;
  if ((rc = Upload2D(ctx, ctx->lut.p, f->sigma_lut_ids, f->sigma_lut_stride,
                     bx, by)))
    return rc;
  if ((rc = Upload2D(ctx, ctx->dcf.p, f->dequant_cf, f->dequant_cf_stride, tx,
                     ty)))
    return rc;
  if ((rc = Upload2D(ctx, ctx->ytox.p, f->ytox_map, f->ytox_stride * 4, tx * 4,
                     ty)))
    return rc;
  if ((rc = Upload2D(ctx, ctx->ytob.p, f->ytob_map, f->ytob_stride * 4, tx * 4,
                     ty)))
    return rc;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->dqmap.p, f->dequant_map, 16 * 256,
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->table.p, f->dequant_table,
                              f->dequant_table_size * 4, hipMemcpyHostToDevice,
                              ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->offsets.p, f->dequant_offsets,
                              f->num_dequant_offsets * 4,
                              hipMemcpyHostToDevice, ctx->stream));
  // The encoder's inverse table (DequantMatrices, pik/quant_weights.cc:
  // 918-923) and whether its quant_table lookups, which index the map by
  // the quant value itself (pik/compressed_image.cc:417, 445), stay inside
  // the 16 x 256 map.
  {
    std::vector<float> inv(f->dequant_table_size);
    for (size_t i = 0; i < inv.size(); ++i) inv[i] = 1.0f / f->dequant_table[i];
    if ((rc = Grow(ctx, &ctx->itable, inv.size() * 4))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->itable.p, inv.data(), inv.size() * 4,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    bool ok = true;
    for (size_t y = 0; y < by && ok; ++y) {
      for (size_t x = 0; x < bx; ++x) {
        const int q = f->quant_field[y * f->quant_field_stride + x];
        const int cf = f->dequant_cf[(y / 8) * f->dequant_cf_stride + x / 8];
        if (cf * 256 + q >= 16 * 256) {
          ok = false;
          break;
        }
      }
    }
    ctx->enc_map_ok = ok;
    ctx->global_scale = int(f->global_scale);
    ctx->quant_dc = int(f->quant_dc);
    for (int c = 0; c < 3; ++c) {
      ctx->dc_matrix0[c] = f->dequant_table[f->dequant_offsets[c]];
    }
  }
  std::vector<int32_t> dot_ids, dot_starts, dot_list;
  if (f->num_dots) {
    HIP_TRY(ctx, hipMemcpyAsync(ctx->dots.p, f->dots,
                                f->num_dots * sizeof(pikgpu_dot),
                                hipMemcpyHostToDevice, ctx->stream));
    BinDots(f, int(bx * 8), int(by * 8), &dot_ids, &dot_starts, &dot_list);
    const size_t nbin = dot_ids.size() + dot_starts.size() + dot_list.size();
    if ((rc = Grow(ctx, &ctx->dotbins, nbin * 4))) return rc;
    int32_t* dst = static_cast<int32_t*>(ctx->dotbins.p);
    HIP_TRY(ctx, hipMemcpyAsync(dst, dot_ids.data(), dot_ids.size() * 4,
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dst + dot_ids.size(), dot_starts.data(),
                                dot_starts.size() * 4, hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dst + dot_ids.size() + dot_starts.size(),
                                dot_list.data(), dot_list.size() * 4,
                                hipMemcpyHostToDevice, ctx->stream));
  }
  std::vector<uint32_t> pred_list;
  int pred_end[4];
  PredClasses(f, &pred_list, pred_end);
  ctx->has_lines = false;
  for (size_t y = 0; y < by && !ctx->has_lines; ++y) {
    for (size_t x = 0; x < bx; ++x) {
      if ((f->ac_strategy[y * f->ac_strategy_stride + x] >> 4) == kLines) {
        ctx->has_lines = true;
        break;
      }
    }
  }
  if ((rc = Grow(ctx, &ctx->predlist, pred_list.size() * 4))) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->predlist.p, pred_list.data(),
                              pred_list.size() * 4, hipMemcpyHostToDevice,
                              ctx->stream));
  // pred2x2 positions that no block writes (ZeroFillImage,
  // pik/compressed_image.cc:976): the same ones on every run of this frame,
  // so once here rather than per run
  HIP_TRY(ctx, hipMemsetAsync(ctx->pred2x2.p, 0, ngroups * 3 * kG2Plane * 4,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));

  FrameBuffers& B = ctx->B;
  memset(&B, 0, sizeof(B));
  B.ac = static_cast<const int16_t*>(ctx->ac.p);
  B.ac_extra = nullptr;
  B.dc16 = static_cast<const int16_t*>(ctx->dc16.p);
  B.qf = static_cast<const int32_t*>(ctx->qf.p);
  B.acs = static_cast<const uint8_t*>(ctx->acs.p);
  B.lut = static_cast<const uint8_t*>(ctx->lut.p);
  B.dcf = static_cast<const uint8_t*>(ctx->dcf.p);
  B.ytox = static_cast<const int32_t*>(ctx->ytox.p);
  B.ytob = static_cast<const int32_t*>(ctx->ytob.p);
  B.dqmap = static_cast<const uint8_t*>(ctx->dqmap.p);
  B.table = static_cast<const float*>(ctx->table.p);
  B.offsets = static_cast<const uint32_t*>(ctx->offsets.p);
  B.num_tables = num_tables;
  B.dots = static_cast<const pikgpu_dot*>(ctx->dots.p);
  B.pred_list = static_cast<const uint32_t*>(ctx->predlist.p);
  for (int k = 0; k < 4; ++k) P.pred_end[k] = pred_end[k];
  P.dot_tiles = int(dot_ids.size());
  P.dot_tiles_x = int(DivCeil(bx * 8, 32));
  if (P.dot_tiles) {
    const int32_t* bins = static_cast<const int32_t*>(ctx->dotbins.p);
    B.dot_tile_ids = bins;
    B.dot_tile_start = bins + dot_ids.size();
    B.dot_list = bins + dot_ids.size() + dot_starts.size();
  }
  B.desc = static_cast<uint4*>(ctx->desc.p);
  B.dcq = static_cast<float*>(ctx->dcq.p);
  B.dc = static_cast<float*>(ctx->dcv.p);
  B.llf = static_cast<float*>(ctx->llf.p);
  B.tmp2x2 = static_cast<float*>(ctx->tmp2x2.p);
  B.lf2x2 = static_cast<float*>(ctx->lf2x2.p);
  B.pred2x2 = static_cast<float*>(ctx->pred2x2.p);
  B.idct = static_cast<float*>(ctx->idct.p);
  B.ar = static_cast<float*>(ctx->ar.p);
  B.out = static_cast<float*>(ctx->out.p);
  B.mm_partial = static_cast<float*>(ctx->mm_partial.p);
  B.mm = static_cast<float*>(ctx->mm.p);
  B.gimg = static_cast<uint8_t*>(ctx->gimg.p);
  B.gpitch = int(gpitch);
  B.grows = int(grows);
  B.ipitch = int(ipitch);
  B.opitch = int(opitch);
  ctx->num_tables = num_tables;
  ctx->uploaded = true;
  return PIKGPU_OK;
}

}  // extern "C"

// Grid of KMinMaxPartial: 256-column strips, row ranges of >= 8 rows, at most
// kMinMaxMaxGroups workgroups (the size of the partials buffer): 16 rows per
// thread at 4096 x 4096, two waves of workgroups on 256 CUs.
static dim3 MinMaxGrid(int w, int h) {
  const int nx = int(DivCeil(w, 256));
  const int ny =
      std::max(1, std::min(int(DivCeil(h, 8)), kMinMaxMaxGroups / nx));
  return dim3(nx, ny);
}
// Grid of KMinMaxPartial4: 1024-column strips (four columns per thread),
// row ranges of >= 8 rows.
static dim3 MinMaxGrid4(int w, int h) {
  const int nx = int(DivCeil(w, 1024));
  const int ny =
      std::max(1, std::min(int(DivCeil(h, 8)), kMinMaxMaxGroups / nx));
  return dim3(nx, ny);
}

static int EnsureStageBuffers(pikgpu_ctx* ctx) {
  const FrameParams& P = ctx->P;
  const size_t nblk = size_t(P.bx) * P.by;
  const size_t img = 3 * size_t(ctx->B.ipitch) * P.py * 4;
  int rc;
  if ((rc = Grow(ctx, &ctx->coeffs, 3 * nblk * 64 * 4))) return rc;
  if ((rc = Grow(ctx, &ctx->guide, img))) return rc;
  if ((rc = Grow(ctx, &ctx->epf, img))) return rc;
  if ((rc = Grow(ctx, &ctx->ar_keep, img))) return rc;
  if ((rc = Grow(ctx, &ctx->gab, img))) return rc;
  ctx->B.coeffs = static_cast<float*>(ctx->coeffs.p);
  ctx->B.guide = static_cast<float*>(ctx->guide.p);
  ctx->B.epf = static_cast<float*>(ctx->epf.p);
  ctx->B.dots_img = static_cast<float*>(ctx->ar_keep.p);
  ctx->B.gab = static_cast<float*>(ctx->gab.p);
  return PIKGPU_OK;
}

// DC: dequant, min/max of the 8-padded DC image, EPF + clamp
// (DequantDC + AdaptiveDCReconstruction).
static void EnqueueDc(const FrameParams& P, const FrameBuffers& B,
                      const DeviceConsts* K, hipStream_t st) {
  const int nblk = P.bx * P.by;
  KDcDequant<<<DivCeil(nblk, 256), 256, 0, st>>>(P, B);
  Img dci;
  dci.p = B.dcq;
  dci.plane = size_t(nblk);
  dci.pitch = P.bx;
  dci.W = int(DivCeil(P.bx, 8) * 8);
  dci.H = int(DivCeil(P.by, 8) * 8);
  dci.vw = P.bx;
  dci.vh = P.by;
  const dim3 mmg = MinMaxGrid(dci.W, dci.H);
  KMinMaxPartial<<<mmg, 256, 0, st>>>(dci, 0, P, B.mm_partial);
  KMinMaxFinal<<<1, 256, 0, st>>>(B.mm_partial, int(mmg.x * mmg.y), B.mm);
  const int grid = XcdGrid(int(DivCeil(dci.W, 32) * DivCeil(dci.H, 32)));
  KEpf<true><<<grid, 256, 0, st>>>(dci, P, B, K, B.mm, B.dc, P.bx,
                                   size_t(nblk));
}

// LF-domain predictions per group: LLF, LF predictions (predict_lf) and,
// with `hf`, the pred2x2 image of the HF prediction.
static void EnqueueLf(const FrameParams& P, const FrameBuffers& B,
                      const DeviceConsts* K, hipStream_t st, bool hf) {
  const size_t ngroups = size_t(P.gx) * P.gy;
  const int cache_pos = (kGroupDimBlocks + 2) * (kGroupDimBlocks + 2);
  const int g2_pos = 4 * cache_pos;
  dim3 gc(DivCeil(cache_pos, 256), ngroups, 3);
  dim3 g2(DivCeil(g2_pos, 256), ngroups, 3);
  // DCT32 / DCT16 first blocks on 8 / 4 lanes (32 / 64 units of block x
  // plane per workgroup), then the cache positions
  const int wg4 = int(DivCeil(3 * (P.pred_end[3] - P.pred_end[2]), 32));
  const int wg2 = int(DivCeil(3 * (P.pred_end[2] - P.pred_end[1]), 64));
  KComputeLlf<<<wg4 + wg2 + gc.x * gc.y * gc.z, 256, 0, st>>>(
      P, B, K, P.predict_lf, wg4, wg2, int(gc.x), int(ngroups));
  if (P.predict_lf) {
    KLfBlur<<<g2, 256, 0, st>>>(P, B);
    KLfFromDc2x2<<<wg4 + wg2 + gc.x * gc.y * gc.z, 256, 0, st>>>(
        P, B, K, wg4, wg2, int(gc.x), int(ngroups));
  }
  if (hf) {
    // (positions of pred2x2 no block writes were zeroed at upload)
    if (P.predict_lf) {
      KPred2x2Border<true><<<gc, 256, 0, st>>>(P, B, K);
    } else {
      KDc2x2Border<<<gc, 256, 0, st>>>(P, B, K);
    }
    if (P.pred_end[3] > 0) {
      // segments as above, then one lane per unit
      const int wg1 = int(DivCeil(3 * P.pred_end[1], 256));
      KPred2x2Interior<false><<<wg4 + wg2 + wg1, 256, 0, st>>>(P, B, K, wg4,
                                                               wg2, nullptr);
    }
  }
}

// DC, LF predictions and the block kernel: everything before the pixel
// filters (the same for full-size and downsampled runs). `ev` (may be
// null) receives events 0..3 around the families.
static hipError_t EnqueuePrefix(const FrameParams& P, const FrameBuffers& B,
                                const DeviceConsts* K, hipStream_t st,
                                const hipEvent_t* ev) {
  hipError_t e = hipSuccess;
#define REC(i) \
  if (ev && (e = hipEventRecord(ev[i], st)) != hipSuccess) return e
  REC(0);
  EnqueueDc(P, B, K, st);
  REC(1);
  EnqueueLf(P, B, K, st, P.predict_hf != 0);
  REC(2);
  {
    const int grid = XcdGrid(int(DivCeil(P.bx, 4) * DivCeil(P.by, 4)));
    KBlocks<<<grid, 256, 0, st>>>(P, B, K);
  }
  REC(3);
#undef REC
  return hipGetLastError();
}

// The full-size reconstruction of a frame whose inputs are in device
// memory (B): DC, predictions, IDCT, EPF + AR (or the plain copy without
// AR), dots, Gaborish and OpsinToLinear into B.out; with B.keep the stage
// copies for pikgpu_stage_dump. Allocates nothing. `ev` (may be null)
// receives events 0..7 between the kernel families.
hipError_t pikgpu::EnqueueFullSize(const FrameParams& P,
                                   const FrameBuffers& B,
                                   const DeviceConsts* K, hipStream_t st,
                                   const hipEvent_t* ev) {
  hipError_t e = EnqueuePrefix(P, B, K, st, ev);
  if (e != hipSuccess) return e;
#define REC(i) \
  if (ev && (e = hipEventRecord(ev[i], st)) != hipSuccess) return e
#define TRY(expr) \
  if ((e = (expr)) != hipSuccess) return e
  const size_t iplane = size_t(B.ipitch) * P.py;
  if (P.ar) {
    Img ii;
    ii.p = B.idct;
    ii.plane = iplane;
    ii.pitch = B.ipitch;
    ii.W = P.px;
    ii.H = P.py;
    ii.vw = P.px;
    ii.vh = P.py;
    const int gab = P.gaborish && !P.use_sharpened;
    // the IDCT image: width a multiple of 8, rows 32-float aligned
    const dim3 mmg = MinMaxGrid4(ii.W, ii.H);
    KMinMaxPartial4<<<mmg, 256, 0, st>>>(ii, gab, P, B.mm_partial);
    KMinMaxFinal<<<1, 256, 0, st>>>(B.mm_partial, int(mmg.x * mmg.y), B.mm);
    const int grid = XcdGrid(int(DivCeil(P.px, 32) * DivCeil(P.py, 32)));
    KGuide<<<grid, 256, 0, st>>>(ii, gab, P, B, B.mm);
    const int ring = 3 * (2 * 6 * (P.px + 12) + 2 * 6 * P.py);
    KGuideRing<<<DivCeil(ring, 256), 256, 0, st>>>(P, B);
    REC(4);
    KEpf<false><<<grid, 256, 0, st>>>(ii, P, B, K, B.mm, B.ar, B.ipitch,
                                      iplane);
  } else {
    REC(4);
    KCopyPlanes<<<DivCeil(3 * iplane, 256), 256, 0, st>>>(B.idct, B.ar,
                                                          3 * iplane);
  }
  REC(5);
  if (B.keep) {
    if (!P.ar) {
      TRY(hipMemcpyAsync(B.guide, B.idct, 3 * iplane * 4,
                         hipMemcpyDeviceToDevice, st));
      TRY(hipMemcpyAsync(B.epf, B.idct, 3 * iplane * 4,
                         hipMemcpyDeviceToDevice, st));
    }
    // B.dots_img keeps the AR output before the dots modify B.ar in place.
    TRY(hipMemcpyAsync(B.dots_img, B.ar, 3 * iplane * 4,
                       hipMemcpyDeviceToDevice, st));
  }
  if (P.dot_tiles) KDots<<<P.dot_tiles, 256, 0, st>>>(P, B);
  REC(6);
  {
    const int grid =
        XcdGrid(int(DivCeil(P.px, 256) * DivCeil(P.py, kFinalRows)));
    KFinal<<<grid, 256, 0, st>>>(P, B, K);
  }
  REC(7);
#undef REC
#undef TRY
  return hipGetLastError();
}

// Enqueues the whole reconstruction of the uploaded frame on stream `st`.
// With `timing`, HIP events between the kernel families record their
// durations (pikgpu_last_timings). Allocates nothing unless stage dumps or
// the 8-bit output need buffers that do not exist yet, so a batch can
// capture it into a graph.
static int Enqueue(pikgpu_ctx* ctx, hipStream_t st, bool keep_stages,
                   bool timing) {
  int rc;
  if (keep_stages && (rc = EnsureStageBuffers(ctx))) return rc;
  FrameParams P = ctx->P;
  FrameBuffers B = ctx->B;
  B.keep = keep_stages != 0;
  P.out_flags = ctx->out_flags;
  P.downsample = ctx->downsample;
  const int ds = P.downsample;
  if (ds != 1) {
    if (keep_stages) {
      return Fail(ctx, PIKGPU_EINVAL, "stage dumps need downsampling factor 1");
    }
    if (P.num_dots != 0) {
      return Fail(ctx, PIKGPU_EUNSUPPORTED,
                  "a frame with dots cannot be downsampled");
    }
  }
  // output size (ShrinkTo of the reference's final stage)
  const int out_xs = int(DivCeil(P.xsize, ds)), out_ys = int(DivCeil(P.ysize, ds));
  if (P.out_flags & kOutSrgb8) {
    B.spitch = int(DivCeil(size_t(out_xs) * 3, 64) * 64);
    if ((rc = Grow(ctx, &ctx->srgb8, size_t(B.spitch) * out_ys))) return rc;
    B.srgb8 = static_cast<uint8_t*>(ctx->srgb8.p);
  }
  const DeviceConsts* K = ctx->consts;
  const hipEvent_t* ev = timing ? ctx->ev : nullptr;

  if (ds != 1) {
    HIP_TRY(ctx, EnqueuePrefix(P, B, K, st, ev));
    // The downsampled branch of FinalizeFrameDecoding: KBlocks wrote the
    // averaged image into B.idct (planes of (py / ds) x ipitch); no AR, no
    // Gaborish, no dots; ShrinkTo and OpsinToLinear in KFinal.
    if (timing) {
      for (int e = 4; e <= 6; ++e) HIP_TRY(ctx, hipEventRecord(ctx->ev[e], st));
    }
    FrameParams Pd = P;
    Pd.px = P.px / ds;
    Pd.py = P.py / ds;
    Pd.xsize = out_xs;
    Pd.ysize = out_ys;
    Pd.gaborish = 0;
    FrameBuffers Bd = B;
    Bd.ar = B.idct;
    const int grid =
        XcdGrid(int(DivCeil(Pd.px, 256) * DivCeil(Pd.py, kFinalRows)));
    KFinal<<<grid, 256, 0, st>>>(Pd, Bd, K);
    if (timing) HIP_TRY(ctx, hipEventRecord(ctx->ev[7], st));
    HIP_TRY(ctx, hipGetLastError());
  } else {
    HIP_TRY(ctx, EnqueueFullSize(P, B, K, st, ev));
  }
  ctx->ran = true;
  ctx->kept = B.keep;
  ctx->ran_flags = P.out_flags;
  ctx->spitch = B.spitch;
  ctx->out_xsize = out_xs;
  ctx->out_ysize = out_ys;
  ctx->ran_downsampled = ds != 1;
  return PIKGPU_OK;
}

extern "C" {

int pikgpu_run(pikgpu_ctx* ctx, int keep_stages) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ctx->timed = true;
  return Enqueue(ctx, ctx->stream, keep_stages != 0, true);
}

int pikgpu_sync(pikgpu_ctx* ctx) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->adj_timed) {
    for (int i = 0; i < 2; ++i) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ctx->adj_ev[i], ctx->adj_ev[i + 1]) ==
          hipSuccess) {
        ctx->timings[kTAdjForward + i] = ms;
      }
    }
    ctx->adj_timed = false;
  }
  if (ctx->ran && ctx->timed) {
    for (int i = 0; i < kNumRunTimers; ++i) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]) == hipSuccess) {
        ctx->timings[i] = ms;
      }
    }
  }
  return PIKGPU_OK;
}

int pikgpu_download(pikgpu_ctx* ctx, float* out_planes[3],
                    size_t out_stride_bytes) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (!ctx->ran) return Fail(ctx, PIKGPU_ESTATE, "pikgpu_run has not run");
  if (!(ctx->ran_flags & kOutLinear)) {
    return Fail(ctx, PIKGPU_ESTATE, "the last run produced no linear output");
  }
  if (out_planes == nullptr || !out_planes[0] || !out_planes[1] ||
      !out_planes[2]) {
    return Fail(ctx, PIKGPU_EINVAL, "null output plane");
  }
  const size_t xs = size_t(ctx->out_xsize), ys = size_t(ctx->out_ysize);
  if (out_stride_bytes < xs * 4) {
    return Fail(ctx, PIKGPU_EINVAL, "output stride smaller than a row");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t oplane = size_t(ctx->B.opitch) * ys;
  const size_t dpitch_b = size_t(ctx->B.opitch) * 4, row_b = xs * 4;
  if (ys > 1 && (dpitch_b != row_b || out_stride_bytes != row_b) &&
      row_b < kStageRowMax) {
    // Short strided rows: chunks of whole device rows (contiguous) to
    // pinned staging, then row by row on the host.
    const size_t per = std::max<size_t>(1, kStageChunk / dpitch_b);
    int rc = StageReserve(ctx, std::min(per, ys) * dpitch_b);
    if (rc != PIKGPU_OK) return rc;
    const uint8_t* p = static_cast<const uint8_t*>(ctx->stage);
    for (int c = 0; c < 3; ++c) {
      uint8_t* q = reinterpret_cast<uint8_t*>(out_planes[c]);
      const float* plane = ctx->B.out + c * oplane;
      for (size_t y0 = 0; y0 < ys; y0 += per) {
        const size_t ny = std::min(per, ys - y0);
        HIP_TRY(ctx, hipMemcpyAsync(ctx->stage, plane + y0 * (dpitch_b / 4),
                                    ny * dpitch_b, hipMemcpyDeviceToHost,
                                    ctx->stream));
        if ((rc = pikgpu_sync(ctx)) != PIKGPU_OK) return rc;
        for (size_t y = 0; y < ny; ++y) {
          memcpy(q + (y0 + y) * out_stride_bytes, p + y * dpitch_b, row_b);
        }
      }
    }
    return PIKGPU_OK;
  }
  for (int c = 0; c < 3; ++c) {
    HIP_TRY(ctx, Copy2D(out_planes[c], out_stride_bytes,
                        ctx->B.out + c * oplane, dpitch_b, row_b, ys,
                        hipMemcpyDeviceToHost, ctx->stream));
  }
  return pikgpu_sync(ctx);
}

int pikgpu_frame_add_pass(pikgpu_ctx* ctx, const int16_t* const ac[3],
                          size_t ac_stride) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  FrameParams& P = ctx->P;
  if (ac == nullptr || !ac[0] || !ac[1] || !ac[2]) {
    return Fail(ctx, PIKGPU_EINVAL, "null AC plane");
  }
  const size_t row = size_t(P.bx) * 64;
  if (ac_stride < row) return Fail(ctx, PIKGPU_EINVAL, "AC stride too small");
  if (P.num_passes >= kMaxPasses) {
    return Fail(ctx, PIKGPU_EINVAL, "too many passes");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t pass_bytes = 3 * row * P.by * sizeof(int16_t);
  const size_t extra = size_t(P.num_passes - 1);
  const size_t need = (extra + 1) * pass_bytes;
  if (ctx->ac_extra.cap < need) {
    // Grow keeping the passes already there. The old buffer is freed only
    // once its contents are in the new one, and the frame's buffer table
    // points at the new one from then on, so a failure at any step leaves a
    // consistent context (the passes so far, in a live buffer).
    DevBuf nb;
    hipError_t e = hipMalloc(&nb.p, need);
    if (e != hipSuccess) {
      return Fail(ctx, PIKGPU_ENOMEM,
                  std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    nb.cap = need;
    if (extra > 0) {
      e = hipMemcpyAsync(nb.p, ctx->ac_extra.p, extra * pass_bytes,
                         hipMemcpyDeviceToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) {
        (void)hipFree(nb.p);
        return Fail(ctx, PIKGPU_EHIP,
                    std::string("copying the earlier passes: ") +
                        hipGetErrorString(e));
      }
    } else {
      // nothing in flight may still read the old buffer
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (ctx->ac_extra.p) (void)hipFree(ctx->ac_extra.p);
    ctx->ac_extra = nb;
    ctx->B.ac_extra = extra > 0 ? static_cast<const int16_t*>(nb.p) : nullptr;
  }
  int16_t* dst = static_cast<int16_t*>(ctx->ac_extra.p) + extra * 3 * row * P.by;
  int rc;
  for (int c = 0; c < 3; ++c) {
    if ((rc = Upload2D(ctx, dst + c * row * P.by, ac[c],
                       ac_stride * sizeof(int16_t), row * sizeof(int16_t),
                       P.by))) {
      return rc;
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  P.num_passes += 1;
  ctx->B.ac_extra = static_cast<const int16_t*>(ctx->ac_extra.p);
  ctx->ran = false;
  return PIKGPU_OK;
}

int pikgpu_decode_frame(pikgpu_ctx* ctx, const uint8_t* data, size_t size,
                        uint32_t xsize, uint32_t ysize, int threads,
                        size_t* consumed) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  char msg[256];
  // the context keeps its parsed frame, so the next frame's host arrays
  // reuse its storage
  int rc = pikgpu_parse_frame(data, size, xsize, ysize, threads, &ctx->parsed,
                              msg, sizeof(msg));
  if (rc != PIKGPU_OK) return Fail(ctx, rc, msg);
  const pikgpu_parsed* parsed = ctx->parsed;
  pikgpu_frame f;
  pikgpu_parsed_info info;
  pikgpu_parsed_get(parsed, &f, &info);
  rc = pikgpu_frame_upload(ctx, &f);
  for (uint32_t i = 1; rc == PIKGPU_OK && i < info.num_passes; ++i) {
    const int16_t* ac[3];
    size_t stride = 0;
    pikgpu_parsed_pass_ac(parsed, i, ac, &stride);
    rc = pikgpu_frame_add_pass(ctx, ac, stride);
  }
  if (rc == PIKGPU_OK) rc = pikgpu_run(ctx, 0);
  if (rc == PIKGPU_OK && consumed != nullptr) *consumed = info.consumed;
  return rc;
}

// Encoder-side quantization of an XYB image under the uploaded frame's side
// information (see pikgpu.h). The quantized AC and DC replace the frame's,
// so pikgpu_run then decodes the round trip.
int pikgpu_quantize(pikgpu_ctx* ctx, const float* const opsin[3],
                    size_t opsin_stride, int16_t* ac_out[3], size_t ac_stride,
                    int16_t* dc_out[3], size_t dc_stride) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  if (opsin == nullptr || !opsin[0] || !opsin[1] || !opsin[2]) {
    return Fail(ctx, PIKGPU_EINVAL, "need three opsin planes");
  }
  FrameParams& P = ctx->P;
  if (opsin_stride < size_t(P.px)) {
    return Fail(ctx, PIKGPU_EINVAL, "opsin stride below the padded width");
  }
  if (P.num_dots != 0 || P.grayscale) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED,
                "quantization of frames with dots or grayscale");
  }
  if (!ctx->enc_map_ok) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED, "quant value past the dequant map");
  }
  const size_t nb = size_t(P.bx) * P.by, row = size_t(P.bx) * 64;
  if ((ac_out && (!ac_out[0] || !ac_out[1] || !ac_out[2] || ac_stride < row)) ||
      (dc_out && (!dc_out[0] || !dc_out[1] || !dc_out[2] ||
                  dc_stride < size_t(P.bx)))) {
    return Fail(ctx, PIKGPU_EINVAL, "bad output planes or strides");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t plane = size_t(P.px) * P.py;
  int rc;
  if ((rc = Grow(ctx, &ctx->eopsin, 3 * plane * 4))) return rc;
  if ((rc = Grow(ctx, &ctx->ecoef, 3 * nb * 64 * 4))) return rc;
  if ((rc = Grow(ctx, &ctx->edcf, 3 * nb * 4))) return rc;
  const bool predict = P.predict_lf || P.predict_hf;
  if (predict && (rc = Grow(ctx, &ctx->ecorner, 3 * nb * 64 * 4))) return rc;
  float* dev_opsin = static_cast<float*>(ctx->eopsin.p);
  for (int c = 0; c < 3; ++c) {
    if ((rc = Upload2D(ctx, dev_opsin + c * plane, opsin[c], opsin_stride * 4,
                       size_t(P.px) * 4, size_t(P.py))))
      return rc;
  }
  // the quantized values replace the frame's own (and any later passes)
  P.num_passes = 1;
  ctx->B.ac_extra = nullptr;
  ctx->ran = false;
  ctx->adj_ready = false;
  const FrameBuffers& B = ctx->B;
  const DeviceConsts* K = ctx->consts;
  hipStream_t st = ctx->stream;
  float* coef = static_cast<float*>(ctx->ecoef.p);
  float* dcf = static_cast<float*>(ctx->edcf.p);
  int16_t* ac16 = static_cast<int16_t*>(ctx->ac.p);
  int16_t* dc16 = static_cast<int16_t*>(ctx->dc16.p);
  const int tiles = int(DivCeil(P.bx, 4) * DivCeil(P.by, 4));
  HIP_TRY(ctx, hipEventRecord(ctx->enc_ev[0], st));
  KEncForward<<<tiles, 256, 0, st>>>(P, B, K, dev_opsin, coef, dcf);
  // Quantizer::QuantizeDC's factors (float arithmetic as the reference's)
  const float gsf = static_cast<float>(ctx->global_scale * (1.0 / 65536));
  const float qdc = gsf * ctx->quant_dc;
  EncDcScale S;
  for (int c = 0; c < 3; ++c) {
    const float inv0 = 1.0f / ctx->dc_matrix0[c];
    S.scale[c] = inv0 * qdc;
  }
  KEncDc<<<DivCeil(nb, 256), 256, 0, st>>>(P, dcf, S, dc16);
  // dc_dec, the LLF and the LF prediction exactly as the decoder computes
  // them (the encoder calls the same ComputeLlf and PredictLf)
  EnqueueDc(P, B, K, st);
  EnqueueLf(P, B, K, st, false);
  const float* itable = static_cast<const float*>(ctx->itable.p);
  if (predict) {
    float* corner = static_cast<float*>(ctx->ecorner.p);
    KEncLf<<<DivCeil(nb, 256), 256, 0, st>>>(P, B, coef, itable, gsf,
                                             corner);
    if (P.predict_hf) {
      // the encoder's pred2x2: the ring as the decoder forms it, the
      // transforms from KEncLf's corners (positions no block writes were
      // zeroed at upload)
      const size_t ngroups = size_t(P.gx) * P.gy;
      const int cache_pos = (kGroupDimBlocks + 2) * (kGroupDimBlocks + 2);
      dim3 gc(DivCeil(cache_pos, 256), ngroups, 3);
      if (P.predict_lf) {
        KPred2x2Border<true><<<gc, 256, 0, st>>>(P, B, K);
      } else {
        KPred2x2Border<false><<<gc, 256, 0, st>>>(P, B, K);
      }
      if (P.pred_end[3] > 0) {
        const int wg4 = int(DivCeil(3 * (P.pred_end[3] - P.pred_end[2]), 32));
        const int wg2 = int(DivCeil(3 * (P.pred_end[2] - P.pred_end[1]), 64));
        const int wg1 = int(DivCeil(3 * P.pred_end[1], 256));
        KPred2x2Interior<true><<<wg4 + wg2 + wg1, 256, 0, st>>>(
            P, B, K, wg4, wg2, corner);
      }
      KEncHfResidual<<<XcdGrid(tiles), 256, 0, st>>>(P, B, coef);
    }
  }
  KEncQuant<<<DivCeil(nb, kEncQuantThreads), kEncQuantThreads, 0, st>>>(
      P, B, coef, itable, gsf, ac16);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->enc_ev[1], st));
  for (int c = 0; c < 3; ++c) {
    if (ac_out) {
      HIP_TRY(ctx, Copy2D(ac_out[c], ac_stride * 2,
                                    ac16 + c * nb * 64, row * 2, row * 2,
                                    size_t(P.by), hipMemcpyDeviceToHost, st));
    }
    if (dc_out) {
      HIP_TRY(ctx, Copy2D(dc_out[c], dc_stride * 2, dc16 + c * nb,
                                    size_t(P.bx) * 2, size_t(P.bx) * 2,
                                    size_t(P.by), hipMemcpyDeviceToHost, st));
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(st));
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ctx->enc_ev[0], ctx->enc_ev[1]) == hipSuccess) {
    ctx->timings[kTEncQuantize] = ms;
  }
  return PIKGPU_OK;
}

int pikgpu_set_downsample(pikgpu_ctx* ctx, int factor) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (factor != 1 && factor != 2 && factor != 4 && factor != 8) {
    return Fail(ctx, PIKGPU_EINVAL, "downsampling factor must be 1, 2, 4 or 8");
  }
  ctx->downsample = factor;
  return PIKGPU_OK;
}

int pikgpu_set_output(pikgpu_ctx* ctx, int outputs) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (outputs <= 0 || (outputs & ~(kOutLinear | kOutSrgb8)) != 0) {
    return Fail(ctx, PIKGPU_EINVAL, "invalid output selection");
  }
  ctx->out_flags = outputs;
  return PIKGPU_OK;
}

int pikgpu_download_srgb8(pikgpu_ctx* ctx, uint8_t* rgb, size_t stride_bytes) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (!ctx->ran) return Fail(ctx, PIKGPU_ESTATE, "pikgpu_run has not run");
  if (!(ctx->ran_flags & kOutSrgb8)) {
    return Fail(ctx, PIKGPU_ESTATE, "the last run produced no 8-bit output");
  }
  const size_t xs = size_t(ctx->out_xsize), ys = size_t(ctx->out_ysize);
  if (rgb == nullptr || stride_bytes < xs * 3) {
    return Fail(ctx, PIKGPU_EINVAL, "bad output buffer or stride");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, Copy2D(rgb, stride_bytes, ctx->srgb8.p,
                                size_t(ctx->spitch), xs * 3, ys,
                                hipMemcpyDeviceToHost, ctx->stream));
  return pikgpu_sync(ctx);
}

int pikgpu_reconstruct(pikgpu_ctx* ctx, const pikgpu_frame* frame,
                       float* out_planes[3], size_t out_stride_bytes) {
  int rc = pikgpu_frame_upload(ctx, frame);
  if (rc) return rc;
  if ((rc = pikgpu_run(ctx, 0))) return rc;
  return pikgpu_download(ctx, out_planes, out_stride_bytes);
}

int pikgpu_stage_size(pikgpu_ctx* ctx, int stage, uint32_t* xsize,
                      uint32_t* ysize) {
  if (ctx == nullptr || xsize == nullptr || ysize == nullptr) {
    return PIKGPU_EINVAL;
  }
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  const FrameParams& P = ctx->P;
  switch (stage) {
    case PIKGPU_STAGE_DC_DEQUANT:
    case PIKGPU_STAGE_DC:
      *xsize = P.bx;
      *ysize = P.by;
      return PIKGPU_OK;
    case PIKGPU_STAGE_COEFFS:
      *xsize = P.bx * 64;
      *ysize = P.by;
      return PIKGPU_OK;
    case PIKGPU_STAGE_LINEAR:
      *xsize = P.xsize;
      *ysize = P.ysize;
      return PIKGPU_OK;
    default:
      if (stage < 0 || stage >= PIKGPU_NUM_STAGES) {
        return Fail(ctx, PIKGPU_EINVAL, "invalid stage");
      }
      *xsize = P.px;
      *ysize = P.py;
      return PIKGPU_OK;
  }
}

int pikgpu_stage_dump(pikgpu_ctx* ctx, int stage, float* planes[3],
                      size_t stride_bytes) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  uint32_t xs, ys;
  int rc = pikgpu_stage_size(ctx, stage, &xs, &ys);
  if (rc) return rc;
  if (ctx->ran && ctx->ran_downsampled && stage != PIKGPU_STAGE_DC_DEQUANT &&
      stage != PIKGPU_STAGE_DC) {
    return Fail(ctx, PIKGPU_ESTATE, "the last run was downsampled");
  }
  if (!ctx->ran || (!ctx->kept && stage != PIKGPU_STAGE_DC_DEQUANT &&
                    stage != PIKGPU_STAGE_DC && stage != PIKGPU_STAGE_IDCT &&
                    stage != PIKGPU_STAGE_AR && stage != PIKGPU_STAGE_LINEAR)) {
    return Fail(ctx, PIKGPU_ESTATE, "stage not kept (run with keep_stages)");
  }
  if (planes == nullptr || !planes[0] || !planes[1] || !planes[2] ||
      stride_bytes < size_t(xs) * 4) {
    return Fail(ctx, PIKGPU_EINVAL, "bad output planes");
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const FrameParams& P = ctx->P;
  const FrameBuffers& B = ctx->B;
  const float* src = nullptr;
  size_t pitch = 0, plane = 0;
  const size_t nblk = size_t(P.bx) * P.by;
  const size_t iplane = size_t(B.ipitch) * P.py;
  switch (stage) {
    case PIKGPU_STAGE_DC_DEQUANT: src = B.dcq; pitch = P.bx; plane = nblk; break;
    case PIKGPU_STAGE_DC: src = B.dc; pitch = P.bx; plane = nblk; break;
    case PIKGPU_STAGE_COEFFS:
      src = B.coeffs; pitch = P.bx * 64; plane = nblk * 64; break;
    case PIKGPU_STAGE_IDCT: src = B.idct; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_GUIDE: src = B.guide; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_EPF: src = B.epf; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_AR:
      // B.ar is modified in place by the dots; the copy taken before them
      // is the AR stage when stages are kept.
      src = ctx->kept ? B.dots_img : B.ar; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_DOTS: src = B.ar; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_GABORISH: src = B.gab; pitch = B.ipitch; plane = iplane; break;
    case PIKGPU_STAGE_LINEAR:
      if (!(ctx->ran_flags & kOutLinear)) {
        return Fail(ctx, PIKGPU_ESTATE, "the last run produced no linear output");
      }
      src = B.out; pitch = B.opitch; plane = size_t(B.opitch) * P.ysize; break;
  }
  for (int c = 0; c < 3; ++c) {
    HIP_TRY(ctx, Copy2D(planes[c], stride_bytes, src + c * plane,
                                  pitch * 4, size_t(xs) * 4, ys,
                                  hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return PIKGPU_OK;
}

int pikgpu_last_timings(pikgpu_ctx* ctx, const char** names, float* ms,
                        int max_entries) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  int n = 0;
  for (int i = 0; i < kNumTimers && n < max_entries; ++i, ++n) {
    if (names) names[n] = kTimerNames[i];
    if (ms) ms[n] = ctx->timings[i];
  }
  return n;
}

// Forward model + adjoint (see pikgpu.h). The LLF / LF caches depend on the
// DC only, so they are computed on the first call after an upload; each
// call then runs KAdjForward and KAdjBackward (and KAdjLoss for the loss).
int pikgpu_fwd_adj(pikgpu_ctx* ctx, const pikgpu_adj_io* io) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (io == nullptr) return Fail(ctx, PIKGPU_EINVAL, "null io");
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  const FrameParams& P0 = ctx->P;
  if (P0.num_passes != 1 || P0.grayscale) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED,
                "the forward model takes single-pass colour frames");
  }
  if (ctx->has_lines) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED,
                "the forward model has no adjoint of the LINES transform");
  }
  const size_t bx = size_t(P0.bx), by = size_t(P0.by);
  const size_t xs = size_t(P0.xsize), ys = size_t(P0.ysize);
  const bool dev = (io->flags & PIKGPU_ADJ_DEVICE_POINTERS) != 0;
  if ((io->flags & ~uint32_t(PIKGPU_ADJ_DEVICE_POINTERS)) != 0) {
    return Fail(ctx, PIKGPU_EINVAL, "unknown flags");
  }
  if (io->qf == nullptr || io->dl_dqf == nullptr) {
    return Fail(ctx, PIKGPU_EINVAL, "qf and dl_dqf are required");
  }
  const bool use_target = io->target[0] != nullptr;
  const float* const* in = use_target ? io->target : io->dl_dlinear;
  if (!in[0] || !in[1] || !in[2]) {
    return Fail(ctx, PIKGPU_EINVAL,
                "need three target planes or three dl_dlinear planes");
  }
  const int nlin = (io->linear[0] != nullptr) + (io->linear[1] != nullptr) +
                   (io->linear[2] != nullptr);
  const int ndc = (io->dl_dcoeffs[0] != nullptr) +
                  (io->dl_dcoeffs[1] != nullptr) +
                  (io->dl_dcoeffs[2] != nullptr);
  if ((nlin != 0 && nlin != 3) || (ndc != 0 && ndc != 3)) {
    return Fail(ctx, PIKGPU_EINVAL,
                "linear and dl_dcoeffs take all three planes or none");
  }
  if (io->loss != nullptr && !use_target) {
    return Fail(ctx, PIKGPU_EINVAL, "the loss needs a target");
  }
  if (io->qf_stride < bx || io->dqf_stride < bx || io->in_stride < xs ||
      (nlin && io->linear_stride < xs) ||
      (ndc && io->dcoeffs_stride < bx * 64)) {
    return Fail(ctx, PIKGPU_EINVAL, "stride smaller than a row");
  }
  if (io->qf_stride > (1u << 30) || io->dqf_stride > (1u << 30) ||
      io->in_stride > (1u << 30) || io->linear_stride > (1u << 30)) {
    return Fail(ctx, PIKGPU_EINVAL, "stride too large");
  }
  if (!dev) {
    for (size_t y = 0; y < by; ++y) {
      const float* r = io->qf + y * io->qf_stride;
      for (size_t x = 0; x < bx; ++x) {
        if (!(r[x] > 0.0f) || !std::isfinite(r[x])) {
          return Fail(ctx, PIKGPU_EINVAL,
                      "quant field values must be finite and positive");
        }
      }
    }
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int tiles = int(DivCeil(bx, 4) * DivCeil(by, 4));
  int rc;
  if ((rc = Grow(ctx, &ctx->adj_loss, size_t(tiles) * 4 + 16))) return rc;

  AdjArgs A;
  memset(&A, 0, sizeof(A));
  const size_t plane_in = xs * ys;
  if (dev) {
    A.qf = io->qf;
    A.qf_pitch = int(io->qf_stride);
    for (int c = 0; c < 3; ++c) {
      (use_target ? A.target : A.upstream)[c] = in[c];
      A.linear[c] = nlin ? io->linear[c] : nullptr;
      A.dcoef[c] = ndc ? io->dl_dcoeffs[c] : nullptr;
    }
    A.in_pitch = int(io->in_stride);
    A.lin_pitch = int(io->linear_stride);
    A.dqf = io->dl_dqf;
    A.dqf_pitch = int(io->dqf_stride);
    A.dc_pitch = io->dcoeffs_stride;
  } else {
    if ((rc = Grow(ctx, &ctx->adj_qf, bx * by * 4)) ||
        (rc = Grow(ctx, &ctx->adj_in, 3 * plane_in * 4)) ||
        (rc = Grow(ctx, &ctx->adj_dqf, bx * by * 4)) ||
        (nlin && (rc = Grow(ctx, &ctx->adj_lin, 3 * plane_in * 4))) ||
        (ndc && (rc = Grow(ctx, &ctx->adj_dc, 3 * by * bx * 64 * 4)))) {
      return rc;
    }
    float* qf = static_cast<float*>(ctx->adj_qf.p);
    float* inb = static_cast<float*>(ctx->adj_in.p);
    if ((rc = Upload2D(ctx, qf, io->qf, io->qf_stride * 4, bx * 4, by))) {
      return rc;
    }
    for (int c = 0; c < 3; ++c) {
      if ((rc = Upload2D(ctx, inb + c * plane_in, in[c], io->in_stride * 4,
                         xs * 4, ys))) {
        return rc;
      }
    }
    A.qf = qf;
    A.qf_pitch = int(bx);
    for (int c = 0; c < 3; ++c) {
      (use_target ? A.target : A.upstream)[c] = inb + c * plane_in;
      A.linear[c] =
          nlin ? static_cast<float*>(ctx->adj_lin.p) + c * plane_in : nullptr;
      A.dcoef[c] =
          ndc ? static_cast<float*>(ctx->adj_dc.p) + c * by * bx * 64 : nullptr;
    }
    A.in_pitch = int(xs);
    A.lin_pitch = int(xs);
    A.dqf = static_cast<float*>(ctx->adj_dqf.p);
    A.dqf_pitch = int(bx);
    A.dc_pitch = bx * 64;
  }
  float* loss_part = static_cast<float*>(ctx->adj_loss.p);
  double* loss_dev = reinterpret_cast<double*>(
      static_cast<char*>(ctx->adj_loss.p) + DivCeil(size_t(tiles) * 4, 8) * 8);
  A.loss_part = use_target ? loss_part : nullptr;

  FrameParams P = ctx->P;
  FrameBuffers B = ctx->B;
  B.keep = 0;
  const DeviceConsts* K = ctx->consts;
  if (!ctx->adj_ready) {
    EnqueueDc(P, B, K, st);
    EnqueueLf(P, B, K, st, false);
    ctx->adj_ready = true;
  }
  const int grid = XcdGrid(tiles);
  HIP_TRY(ctx, hipEventRecord(ctx->adj_ev[0], st));
  KAdjForward<<<grid, 256, 0, st>>>(P, B, K, A);
  HIP_TRY(ctx, hipEventRecord(ctx->adj_ev[1], st));
  KAdjBackward<<<grid, 256, 0, st>>>(P, B, K, A);
  HIP_TRY(ctx, hipEventRecord(ctx->adj_ev[2], st));
  if (io->loss) KAdjLoss<<<1, 256, 0, st>>>(loss_part, tiles, loss_dev);
  HIP_TRY(ctx, hipGetLastError());
  ctx->adj_timed = true;
  // pikgpu_run's output is not what the device buffers hold any more.
  ctx->ran = false;
  if (dev) {
    if (io->loss) {
      HIP_TRY(ctx, hipMemcpyAsync(io->loss, loss_dev, 8,
                                  hipMemcpyDeviceToDevice, st));
    }
    return PIKGPU_OK;
  }
  HIP_TRY(ctx, Copy2D(io->dl_dqf, io->dqf_stride * 4, A.dqf, bx * 4,
                                bx * 4, by, hipMemcpyDeviceToHost, st));
  for (int c = 0; c < 3; ++c) {
    if (nlin) {
      HIP_TRY(ctx, Copy2D(io->linear[c], io->linear_stride * 4,
                                    A.linear[c], xs * 4, xs * 4, ys,
                                    hipMemcpyDeviceToHost, st));
    }
    if (ndc) {
      HIP_TRY(ctx, Copy2D(io->dl_dcoeffs[c], io->dcoeffs_stride * 4,
                                    A.dcoef[c], bx * 64 * 4, bx * 64 * 4, by,
                                    hipMemcpyDeviceToHost, st));
    }
  }
  if (io->loss) {
    HIP_TRY(ctx, hipMemcpyAsync(io->loss, loss_dev, 8, hipMemcpyDeviceToHost,
                                st));
  }
  return pikgpu_sync(ctx);
}


// Gradient quant-field search (see pikgpu.h): device-resident loop of
// quantization (KEncQuant), forward + adjoint (KAdjForward / KAdjBackward),
// distortion attribution (KSearchDist) and Adam (KSearchUpdate).
int pikgpu_search_quant_field(pikgpu_ctx* ctx,
                              const pikgpu_search_params* prm,
                              const pikgpu_search_io* io) {
  if (ctx == nullptr) return PIKGPU_EINVAL;
  if (prm == nullptr || io == nullptr) {
    return Fail(ctx, PIKGPU_EINVAL, "null params or io");
  }
  if (!ctx->uploaded) return Fail(ctx, PIKGPU_ESTATE, "no frame uploaded");
  FrameParams& P = ctx->P;
  if (P.predict_lf || P.predict_hf || P.num_dots != 0 || P.grayscale) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED,
                "search of frames with predictions, dots or grayscale");
  }
  if (ctx->has_lines) {
    return Fail(ctx, PIKGPU_EUNSUPPORTED,
                "the search has no adjoint of the LINES transform");
  }
  if (prm->steps < 0 || prm->steps > 1000 || !(prm->learning_rate > 0.0f) ||
      !(prm->beta1 >= 0.0f && prm->beta1 < 1.0f) ||
      !(prm->beta2 >= 0.0f && prm->beta2 < 1.0f) || !(prm->eps > 0.0f) ||
      prm->qf_max < 1 || prm->qf_max > 256 || !std::isfinite(prm->lambda) ||
      !(prm->trust_low > 0.0f && prm->trust_low <= 1.0f) ||
      !(prm->trust_high >= 1.0f && std::isfinite(prm->trust_high))) {
    return Fail(ctx, PIKGPU_EINVAL, "invalid search parameters");
  }
  if (!io->opsin[0] || !io->opsin[1] || !io->opsin[2] || !io->target[0] ||
      !io->target[1] || !io->target[2]) {
    return Fail(ctx, PIKGPU_EINVAL, "need three opsin and three target planes");
  }
  const size_t bx = size_t(P.bx), by = size_t(P.by), nb = bx * by;
  const size_t xs = size_t(P.xsize), ys = size_t(P.ysize);
  const int nac = (io->ac[0] != nullptr) + (io->ac[1] != nullptr) +
                  (io->ac[2] != nullptr);
  const int ndc = (io->dc[0] != nullptr) + (io->dc[1] != nullptr) +
                  (io->dc[2] != nullptr);
  if ((nac != 0 && nac != 3) || (ndc != 0 && ndc != 3)) {
    return Fail(ctx, PIKGPU_EINVAL, "ac and dc outputs take all three planes");
  }
  if (io->opsin_stride < size_t(P.px) || io->target_stride < xs ||
      (io->weight && io->weight_stride < bx) ||
      (io->quant_field && io->quant_field_stride < bx) ||
      (nac && io->ac_stride < bx * 64) || (ndc && io->dc_stride < bx)) {
    return Fail(ctx, PIKGPU_EINVAL, "stride smaller than a row");
  }
  // the encoder's dequant-map lookup (indexed by the quant value itself)
  // must stay inside the map for every value the search may take
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int qmax = prm->qf_max;
  {
    std::vector<uint8_t> cf(size_t(P.tx) * P.ty);
    HIP_TRY(ctx, hipMemcpy(cf.data(), ctx->B.dcf, cf.size(),
                           hipMemcpyDeviceToHost));
    for (uint8_t v : cf) {
      if (v == 15) qmax = std::min(qmax, 255);
    }
  }
  hipStream_t st = ctx->stream;
  const size_t plane = size_t(P.px) * P.py;
  const size_t tplane = xs * ys;
  const int steps = prm->steps;
  const int round_groups = int(DivCeil(nb, 256));
  const int tiles = int(DivCeil(bx, 4) * DivCeil(by, 4));
  // search state: u, m, v, weight, dist, u0 (nb floats each), lambda, rate
  // partials (round_groups floats), loss / rate histories (steps + 1 doubles)
  const size_t hist_off = DivCeil((6 * nb + 1 + size_t(round_groups)) * 4, 8) * 8;
  int rc;
  if ((rc = Grow(ctx, &ctx->eopsin, 3 * plane * 4)) ||
      (rc = Grow(ctx, &ctx->ecoef, 3 * nb * 64 * 4)) ||
      (rc = Grow(ctx, &ctx->edcf, 3 * nb * 4)) ||
      (rc = Grow(ctx, &ctx->adj_qf, nb * 4)) ||
      (rc = Grow(ctx, &ctx->adj_in, 3 * tplane * 4)) ||
      (rc = Grow(ctx, &ctx->adj_dqf, nb * 4)) ||
      (rc = Grow(ctx, &ctx->adj_dc, 3 * nb * 64 * 4)) ||
      (rc = Grow(ctx, &ctx->adj_loss, size_t(tiles) * 4 + 16)) ||
      (rc = Grow(ctx, &ctx->srch, hist_off + 2 * size_t(steps + 1) * 8))) {
    return rc;
  }
  float* dev_opsin = static_cast<float*>(ctx->eopsin.p);
  float* target = static_cast<float*>(ctx->adj_in.p);
  for (int c = 0; c < 3; ++c) {
    if ((rc = Upload2D(ctx, dev_opsin + c * plane, io->opsin[c],
                       io->opsin_stride * 4, size_t(P.px) * 4, size_t(P.py))) ||
        (rc = Upload2D(ctx, target + c * tplane, io->target[c],
                       io->target_stride * 4, xs * 4, ys))) {
      return rc;
    }
  }
  float* st_u = static_cast<float*>(ctx->srch.p);
  float* st_m = st_u + nb;
  float* st_v = st_m + nb;
  float* st_w = st_v + nb;
  float* st_d = st_w + nb;
  float* st_u0 = st_d + nb;
  float* st_lam = st_u0 + nb;
  float* st_rate = st_lam + 1;
  double* h_loss = reinterpret_cast<double*>(
      static_cast<char*>(ctx->srch.p) + hist_off);
  double* h_rate = h_loss + (steps + 1);
  if (io->weight) {
    if ((rc = Upload2D(ctx, st_w, io->weight, io->weight_stride * 4, bx * 4,
                       by))) {
      return rc;
    }
  }
  // the quantized values replace the frame's own (and any later passes)
  P.num_passes = 1;
  ctx->B.ac_extra = nullptr;
  ctx->ran = false;
  FrameBuffers& B = ctx->B;
  const DeviceConsts* K = ctx->consts;
  float* coef = static_cast<float*>(ctx->ecoef.p);
  float* dcf = static_cast<float*>(ctx->edcf.p);
  int16_t* ac16 = static_cast<int16_t*>(ctx->ac.p);
  int16_t* dc16 = static_cast<int16_t*>(ctx->dc16.p);
  int32_t* qf = static_cast<int32_t*>(ctx->qf.p);
  float* qfr = static_cast<float*>(ctx->adj_qf.p);
  float* grad = static_cast<float*>(ctx->adj_dc.p);
  // forward transforms and the DC: independent of the AC field
  KEncForward<<<tiles, 256, 0, st>>>(P, B, K, dev_opsin, coef, dcf);
  const float gsf = static_cast<float>(ctx->global_scale * (1.0 / 65536));
  {
    EncDcScale S;
    const float qdc = gsf * ctx->quant_dc;
    for (int c = 0; c < 3; ++c) S.scale[c] = (1.0f / ctx->dc_matrix0[c]) * qdc;
    KEncDc<<<DivCeil(nb, 256), 256, 0, st>>>(P, dcf, S, dc16);
  }
  EnqueueDc(P, B, K, st);
  EnqueueLf(P, B, K, st, false);  // the LLF (CopyLlf) and the LF caches
  ctx->adj_ready = true;
  KSearchInit<<<DivCeil(nb, 256), 256, 0, st>>>(P, B, st_u0, st_u, st_m,
                                                  st_v);

  AdjArgs A;
  memset(&A, 0, sizeof(A));
  A.qf = qfr;
  A.qf_pitch = int(bx);
  for (int c = 0; c < 3; ++c) {
    A.target[c] = target + c * tplane;
    A.dcoef[c] = grad + c * nb * 64;
  }
  A.in_pitch = int(xs);
  A.dqf = static_cast<float*>(ctx->adj_dqf.p);
  A.dqf_pitch = int(bx);
  A.dc_pitch = bx * 64;
  A.loss_part = static_cast<float*>(ctx->adj_loss.p);
  SearchArgs S;
  S.weight = io->weight ? st_w : nullptr;
  S.lambda = st_lam;
  S.lr = prm->learning_rate;
  S.beta1 = prm->beta1;
  S.beta2 = prm->beta2;
  S.eps = prm->eps;
  S.ulo = 0.0f;
  S.uhi = std::log(float(qmax));
  S.dlo = std::log(prm->trust_low);
  S.dhi = std::log(prm->trust_high);
  const float* itable = static_cast<const float*>(ctx->itable.p);
  const int grid = XcdGrid(tiles);
  for (int t = 0; t <= steps; ++t) {
    KSearchRound<<<round_groups, 256, 0, st>>>(P, B, st_u, qmax, qf, qfr,
                                               st_rate);
    KAdjLoss<<<1, 256, 0, st>>>(st_rate, round_groups, h_rate + t);
    KEncQuant<<<DivCeil(nb, kEncQuantThreads), kEncQuantThreads, 0, st>>>(
        P, B, coef, itable, gsf, ac16);
    KAdjForward<<<grid, 256, 0, st>>>(P, B, K, A);
    KAdjBackward<<<grid, 256, 0, st>>>(P, B, K, A);
    KAdjLoss<<<1, 256, 0, st>>>(A.loss_part, tiles, h_loss + t);
    if (t == steps) break;
    KSearchDist<<<DivCeil(nb, 4), 256, 0, st>>>(P, B, coef, grad, st_d);
    if (t == 0) {
      KSearchLambda<<<1, 256, 0, st>>>(
          st_d, S.weight, int(nb), prm->lambda,
          prm->lambda < 0.0f ? -prm->lambda : 1.0f, st_lam);
    }
    KSearchUpdate<<<DivCeil(nb, 256), 256, 0, st>>>(P, B, S, t + 1, st_d,
                                                    st_u0, st_u, st_m, st_v);
  }
  HIP_TRY(ctx, hipGetLastError());
  // outputs
  if (io->quant_field) {
    HIP_TRY(ctx, Copy2D(io->quant_field, io->quant_field_stride * 4,
                                  qf, bx * 4, bx * 4, by,
                                  hipMemcpyDeviceToHost, st));
  }
  for (int c = 0; c < 3; ++c) {
    if (nac) {
      HIP_TRY(ctx, Copy2D(io->ac[c], io->ac_stride * 2,
                                    ac16 + c * nb * 64, bx * 128, bx * 128, by,
                                    hipMemcpyDeviceToHost, st));
    }
    if (ndc) {
      HIP_TRY(ctx, Copy2D(io->dc[c], io->dc_stride * 2,
                                    dc16 + c * nb, bx * 2, bx * 2, by,
                                    hipMemcpyDeviceToHost, st));
    }
  }
  if (io->loss) {
    HIP_TRY(ctx, hipMemcpyAsync(io->loss, h_loss, size_t(steps + 1) * 8,
                                hipMemcpyDeviceToHost, st));
  }
  if (io->rate) {
    HIP_TRY(ctx, hipMemcpyAsync(io->rate, h_rate, size_t(steps + 1) * 8,
                                hipMemcpyDeviceToHost, st));
  }
  if (io->lambda_used) {
    HIP_TRY(ctx, hipMemcpyAsync(io->lambda_used, st_lam, 4,
                                hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return PIKGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: several frames of any sizes reconstructed by one graph launch.
// Each frame keeps its own buffers (a context per frame). In the queue
// schedule the graph holds one launch per phase of the reconstruction over
// all frames: the batched kernels (kernels.h, BatchSpan) find their frame
// and its workgroup from a span table, so every phase fills the GPU however
// small the frames are. In the branch schedule the launch sequences of all
// frames are captured into kBranches parallel branches of the graph.

namespace {
constexpr int kBranches = 32;

typedef void (*BatchKernel)(const BatchFrame*, const BatchSpan*,
                            const uint32_t*, uint32_t, const DeviceConsts*);

// The phases in launch order: name, kernel, whether it can loop
// (persistent), and its minimum workgroups per CU for the loop's grid.
struct BatchPhaseDef {
  const char* name;
  BatchKernel kernel;
  bool loop;
};
const BatchPhaseDef kBatchPhases[PIKGPU_BATCH_PHASES] = {
    {"dc_dequant", KDcDequantBatch, true},
    {"dc_minmax", KMinMaxPartialDcBatch, true},
    {"dc_minmax_final", KMinMaxFinalBatch, true},
    {"dc_epf", KEpfDcBatch, false},
    {"llf", KComputeLlfBatch, true},
    {"lf_blur", KLfBlurBatch, true},
    {"lf_from_dc2x2", KLfFromDc2x2Batch, true},
    {"pred2x2_border", KPred2x2BorderBatch, true},
    {"dc2x2_border", KDc2x2BorderBatch, true},
    {"pred2x2_interior", KPred2x2InteriorBatch, false},
    {"blocks", KBlocksBatch, false},
    {"minmax", KMinMaxPartial4Batch, true},
    {"minmax_final", KMinMaxFinalBatch, true},
    {"guide", KGuideBatch, false},
    {"guide_ring", KGuideRingBatch, true},
    {"epf", KEpfBatch, false},
    {"copy_planes", KCopyPlanesBatch, true},
    {"dots", KDotsBatch, true},
    {"final", KFinalBatch, true},
};
}  // namespace

struct pikgpu_batch {
  int device = 0;
  pikgpu_config config = {PIKGPU_ABI_VERSION, 0};
  std::string error;
  std::vector<pikgpu_ctx*> items;
  int n = 0;
  hipStream_t main = nullptr;
  hipStream_t branch[kBranches] = {};
  hipEvent_t fork = nullptr, join[kBranches] = {}, t0 = nullptr, t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ran = false;
  float last_ms = 0.0f;
  pikgpu_batch_schedule sched = {PIKGPU_BATCH_QUEUE, 1, {}};
  int num_cus = 256;
  // queue schedule: the frames' parameters, the phases' span tables and
  // their workgroup -> span maps
  DevBuf frames, spans, maps;
};

namespace {

int BFail(pikgpu_batch* b, int code, const std::string& msg) {
  if (b) b->error = msg;
  return code;
}

#define BHIP_TRY(b, expr)                                                   \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      return BFail(b, PIKGPU_EHIP,                                          \
                   std::string(#expr) + ": " + hipGetErrorString(e_));      \
    }                                                                       \
  } while (0)

int GrowRaw(DevBuf* b, size_t bytes) {
  if (b->cap >= bytes) return PIKGPU_OK;
  if (b->p) (void)hipFree(b->p);
  b->p = nullptr;
  b->cap = 0;
  if (hipMalloc(&b->p, bytes) != hipSuccess) return PIKGPU_ENOMEM;
  b->cap = bytes;
  return PIKGPU_OK;
}

void DropGraph(pikgpu_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// One frame's grid and arguments of phase p, as its plain launch in
// Enqueue has them; false when the phase does not run for the frame.
bool PhaseGrid(int p, const FrameParams& P, const FrameBuffers& B, int* nx,
               int* ny, int* nz, int* arg) {
  *nx = *ny = *nz = 1;
  const int nblk = P.bx * P.by;
  const size_t ngroups = size_t(P.gx) * P.gy;
  const int cache_pos = (kGroupDimBlocks + 2) * (kGroupDimBlocks + 2);
  const int gcx = int(DivCeil(cache_pos, 256));
  const int g2x = int(DivCeil(4 * cache_pos, 256));
  const int wg4 = int(DivCeil(3 * (P.pred_end[3] - P.pred_end[2]), 32));
  const int wg2 = int(DivCeil(3 * (P.pred_end[2] - P.pred_end[1]), 64));
  const Img dci = DcImage(P, B);
  const Img ii = IdctImage(P, B);
  const bool hf = P.predict_hf != 0;
  switch (p) {
    case 0:  // dc_dequant
      *nx = int(DivCeil(nblk, 256));
      return true;
    case 1: {  // dc_minmax
      const dim3 g = MinMaxGrid(dci.W, dci.H);
      *nx = int(g.x);
      *ny = int(g.y);
      return true;
    }
    case 2: {  // dc_minmax_final
      const dim3 g = MinMaxGrid(dci.W, dci.H);
      arg[0] = int(g.x * g.y);
      return true;
    }
    case 3:  // dc_epf
      *nx = XcdGrid(int(DivCeil(dci.W, 32) * DivCeil(dci.H, 32)));
      return true;
    case 4:  // llf
      *nx = wg4 + wg2 + gcx * int(ngroups) * 3;
      arg[0] = P.predict_lf;
      arg[1] = wg4;
      arg[2] = wg2;
      arg[3] = gcx;
      arg[4] = int(ngroups);
      return true;
    case 5:  // lf_blur
      *nx = g2x;
      *ny = int(ngroups);
      *nz = 3;
      return P.predict_lf != 0;
    case 6:  // lf_from_dc2x2
      *nx = wg4 + wg2 + gcx * int(ngroups) * 3;
      arg[0] = wg4;
      arg[1] = wg2;
      arg[2] = gcx;
      arg[3] = int(ngroups);
      return P.predict_lf != 0;
    case 7:  // pred2x2_border
    case 8:  // dc2x2_border
      *nx = gcx;
      *ny = int(ngroups);
      *nz = 3;
      return hf && (P.predict_lf != 0) == (p == 7);
    case 9: {  // pred2x2_interior
      const int wg1 = int(DivCeil(3 * P.pred_end[1], 256));
      *nx = wg4 + wg2 + wg1;
      arg[0] = wg4;
      arg[1] = wg2;
      return hf && P.pred_end[3] > 0;
    }
    case 10:  // blocks
      *nx = XcdGrid(int(DivCeil(P.bx, 4) * DivCeil(P.by, 4)));
      return true;
    case 11: {  // minmax
      const dim3 g = MinMaxGrid4(ii.W, ii.H);
      *nx = int(g.x);
      *ny = int(g.y);
      arg[0] = P.gaborish && !P.use_sharpened;
      return P.ar != 0;
    }
    case 12: {  // minmax_final
      const dim3 g = MinMaxGrid4(ii.W, ii.H);
      arg[0] = int(g.x * g.y);
      return P.ar != 0;
    }
    case 13:  // guide
      *nx = XcdGrid(int(DivCeil(P.px, 32) * DivCeil(P.py, 32)));
      arg[0] = P.gaborish && !P.use_sharpened;
      return P.ar != 0;
    case 14:  // guide_ring
      *nx = int(DivCeil(3 * (2 * 6 * (P.px + 12) + 2 * 6 * P.py), 256));
      return P.ar != 0;
    case 15:  // epf
      *nx = XcdGrid(int(DivCeil(P.px, 32) * DivCeil(P.py, 32)));
      return P.ar != 0;
    case 16:  // copy_planes
      *nx = int(DivCeil(3 * size_t(B.ipitch) * P.py, 256));
      return P.ar == 0;
    case 17:  // dots
      *nx = P.dot_tiles;
      return P.dot_tiles > 0;
    case 18:  // final
      *nx = XcdGrid(int(DivCeil(P.px, 256) * DivCeil(P.py, kFinalRows)));
      return true;
  }
  return false;
}

// The queue schedule's graph: the frames' parameters and span tables in
// device memory, then one launch per phase.
int BuildQueueGraph(pikgpu_batch* b) {
  std::vector<BatchFrame> frames(b->n);
  std::vector<int> order(b->n);
  for (int i = 0; i < b->n; ++i) {
    pikgpu_ctx* c = b->items[i];
    frames[i].P = c->P;
    frames[i].P.out_flags = kOutLinear;
    frames[i].P.downsample = 1;
    frames[i].B = c->B;
    frames[i].B.keep = 0;
    order[i] = i;
  }
  if (b->sched.order != 0) {
    const bool big_first = b->sched.order == 1;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const long ax = long(frames[x].P.px) * frames[x].P.py;
      const long ay = long(frames[y].P.px) * frames[y].P.py;
      return big_first ? ax > ay : ax < ay;
    });
  }
  struct Launch {
    int phase;
    size_t first_span;
    int ns;
    uint32_t total;
    size_t first_map;
  };
  std::vector<BatchSpan> spans;
  std::vector<uint32_t> maps;
  std::vector<Launch> launches;
  for (int p = 0; p < PIKGPU_BATCH_PHASES; ++p) {
    Launch L = {p, spans.size(), 0, 0, maps.size()};
    uint32_t at = 0;
    for (int i : order) {
      BatchSpan sp;
      memset(&sp, 0, sizeof(sp));
      if (!PhaseGrid(p, frames[i].P, frames[i].B, &sp.nx, &sp.ny, &sp.nz,
                     sp.arg)) {
        continue;
      }
      sp.frame = uint32_t(i);
      sp.first = at;
      sp.count = uint32_t(sp.nx) * sp.ny * sp.nz;
      if (sp.count == 0) continue;
      at = uint32_t(XcdGrid(int(at + sp.count)));
      spans.push_back(sp);
      L.ns += 1;
    }
    L.total = at;
    if (L.ns == 0) continue;
    maps.resize(L.first_map + at, kNoSpan);
    for (int j = 0; j < L.ns; ++j) {
      const BatchSpan& sp = spans[L.first_span + j];
      std::fill_n(maps.begin() + L.first_map + sp.first, sp.count,
                  uint32_t(j));
    }
    launches.push_back(L);
  }
  int rc;
  if ((rc = GrowRaw(&b->frames, frames.size() * sizeof(BatchFrame))) ||
      (rc = GrowRaw(&b->spans, std::max<size_t>(1, spans.size()) *
                                   sizeof(BatchSpan))) ||
      (rc = GrowRaw(&b->maps,
                    std::max<size_t>(1, maps.size()) * sizeof(uint32_t)))) {
    return BFail(b, rc, "batch tables: out of device memory");
  }
  BHIP_TRY(b, hipMemcpy(b->frames.p, frames.data(),
                        frames.size() * sizeof(BatchFrame),
                        hipMemcpyHostToDevice));
  if (!spans.empty()) {
    BHIP_TRY(b, hipMemcpy(b->spans.p, spans.data(),
                          spans.size() * sizeof(BatchSpan),
                          hipMemcpyHostToDevice));
    BHIP_TRY(b, hipMemcpy(b->maps.p, maps.data(),
                          maps.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
  }
  const BatchFrame* F = static_cast<const BatchFrame*>(b->frames.p);
  const BatchSpan* S0 = static_cast<const BatchSpan*>(b->spans.p);
  const uint32_t* M0 = static_cast<const uint32_t*>(b->maps.p);
  const DeviceConsts* K = b->items[0]->consts;
  BHIP_TRY(b, hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (const Launch& L : launches) {
    if (e != hipSuccess) break;
    const BatchPhaseDef& d = kBatchPhases[L.phase];
    uint32_t grid = L.total;
    const int k = b->sched.persist[L.phase];
    if (d.loop && k > 0) {
      grid = std::min<uint32_t>(grid, uint32_t(XcdGrid(k * b->num_cus)));
    }
    const BatchSpan* S = S0 + L.first_span;
    const uint32_t* M = M0 + L.first_map;
    uint32_t total = L.total;
    void* args[] = {&F, &S, &M, &total, &K};
    e = hipLaunchKernel(reinterpret_cast<const void*>(d.kernel), dim3(grid),
                        dim3(256), args, 0, b->main);
  }
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(b->main, &g);
  if (e != hipSuccess || e_end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return BFail(b, PIKGPU_EHIP,
                 std::string("graph capture: ") +
                     hipGetErrorString(e != hipSuccess ? e : e_end));
  }
  b->graph = g;
  BHIP_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  // what pikgpu_download reads of each frame's run
  for (int i = 0; i < b->n; ++i) {
    pikgpu_ctx* c = b->items[i];
    c->ran = true;
    c->kept = false;
    c->ran_flags = kOutLinear;
    c->out_xsize = int(c->P.xsize);
    c->out_ysize = int(c->P.ysize);
    c->ran_downsampled = false;
    c->timed = false;
  }
  return PIKGPU_OK;
}

int BuildGraph(pikgpu_batch* b) {
  DropGraph(b);
  if (b->sched.mode == PIKGPU_BATCH_QUEUE) return BuildQueueGraph(b);
  BHIP_TRY(b, hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
  int rc = PIKGPU_OK;
  std::string err;
  hipError_t e = hipEventRecord(b->fork, b->main);
  for (int k = 0; k < kBranches && e == hipSuccess; ++k) {
    e = hipStreamWaitEvent(b->branch[k], b->fork, 0);
  }
  for (int i = 0; i < b->n && e == hipSuccess && rc == PIKGPU_OK; ++i) {
    rc = Enqueue(b->items[i], b->branch[i % kBranches], false, false);
    if (rc != PIKGPU_OK) {
      err = "frame " + std::to_string(i) + ": " + b->items[i]->error;
    }
  }
  for (int k = 0; k < kBranches && e == hipSuccess; ++k) {
    e = hipEventRecord(b->join[k], b->branch[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->main, b->join[k], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(b->main, &g);
  if (rc != PIKGPU_OK || e != hipSuccess || e_end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != PIKGPU_OK) return BFail(b, rc, err);
    return BFail(b, PIKGPU_EHIP,
                 std::string("graph capture: ") +
                     hipGetErrorString(e != hipSuccess ? e : e_end));
  }
  b->graph = g;
  BHIP_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return PIKGPU_OK;
}

}  // namespace

pikgpu_batch* pikgpu_batch_create(int device, const pikgpu_config* config) {
  if (config && config->abi_version != PIKGPU_ABI_VERSION) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  pikgpu_batch* b = new pikgpu_batch;
  b->device = device;
  if (config) b->config = *config;
  bool ok = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreate(&b->t0) == hipSuccess &&
            hipEventCreate(&b->t1) == hipSuccess;
  for (int k = 0; k < kBranches && ok; ++k) {
    ok = hipStreamCreateWithFlags(&b->branch[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    pikgpu_batch_destroy(b);
    return nullptr;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            device) == hipSuccess && cus > 0) {
    b->num_cus = cus;
  }
  return b;
}

void pikgpu_batch_destroy(pikgpu_batch* b) {
  if (b == nullptr) return;
  (void)hipSetDevice(b->device);
  if (b->main) (void)hipStreamSynchronize(b->main);
  DropGraph(b);
  for (pikgpu_ctx* c : b->items) pikgpu_destroy(c);
  for (int k = 0; k < kBranches; ++k) {
    if (b->branch[k]) (void)hipStreamDestroy(b->branch[k]);
    if (b->join[k]) (void)hipEventDestroy(b->join[k]);
  }
  if (b->frames.p) (void)hipFree(b->frames.p);
  if (b->spans.p) (void)hipFree(b->spans.p);
  if (b->maps.p) (void)hipFree(b->maps.p);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->main) (void)hipStreamDestroy(b->main);
  delete b;
}

const char* pikgpu_batch_last_error(const pikgpu_batch* b) {
  return b ? b->error.c_str() : "null batch";
}

int pikgpu_batch_upload(pikgpu_batch* b, const pikgpu_frame* frames, int n) {
  if (b == nullptr) return PIKGPU_EINVAL;
  if (n < 1 || frames == nullptr) return BFail(b, PIKGPU_EINVAL, "empty batch");
  BHIP_TRY(b, hipSetDevice(b->device));
  BHIP_TRY(b, hipStreamSynchronize(b->main));
  DropGraph(b);
  b->ran = false;
  b->n = 0;
  while (int(b->items.size()) < n) {
    pikgpu_ctx* c = pikgpu_create(b->device, &b->config);
    if (c == nullptr) return BFail(b, PIKGPU_EHIP, "context creation failed");
    b->items.push_back(c);
  }
  for (int i = 0; i < n; ++i) {
    const int rc = pikgpu_frame_upload(b->items[i], &frames[i]);
    if (rc != PIKGPU_OK) {
      return BFail(b, rc, "frame " + std::to_string(i) + ": " +
                              b->items[i]->error);
    }
  }
  b->n = n;
  return PIKGPU_OK;
}

int pikgpu_batch_add_pass(pikgpu_batch* b, int i, const int16_t* const ac[3],
                          size_t ac_stride) {
  if (b == nullptr) return PIKGPU_EINVAL;
  if (i < 0 || i >= b->n) return BFail(b, PIKGPU_EINVAL, "no such frame");
  // a running graph may read the frame's pass buffer, which can move
  BHIP_TRY(b, hipSetDevice(b->device));
  BHIP_TRY(b, hipStreamSynchronize(b->main));
  const int rc = pikgpu_frame_add_pass(b->items[i], ac, ac_stride);
  if (rc != PIKGPU_OK) {
    return BFail(b, rc, "frame " + std::to_string(i) + ": " +
                            pikgpu_last_error(b->items[i]));
  }
  // the frame table (pass count, pass buffer) changed: new graph
  DropGraph(b);
  b->ran = false;
  return PIKGPU_OK;
}

int pikgpu_batch_set_schedule(pikgpu_batch* b,
                              const pikgpu_batch_schedule* s) {
  if (b == nullptr) return PIKGPU_EINVAL;
  if (s == nullptr || (s->mode != PIKGPU_BATCH_QUEUE &&
                       s->mode != PIKGPU_BATCH_BRANCHES) ||
      s->order < 0 || s->order > 2) {
    return BFail(b, PIKGPU_EINVAL, "invalid schedule");
  }
  for (int p = 0; p < PIKGPU_BATCH_PHASES; ++p) {
    if (s->persist[p] < 0 || s->persist[p] > 16) {
      return BFail(b, PIKGPU_EINVAL, "persist outside 0 .. 16");
    }
  }
  BHIP_TRY(b, hipSetDevice(b->device));
  BHIP_TRY(b, hipStreamSynchronize(b->main));
  DropGraph(b);
  b->sched = *s;
  return PIKGPU_OK;
}

int pikgpu_batch_get_schedule(const pikgpu_batch* b, pikgpu_batch_schedule* s) {
  if (b == nullptr || s == nullptr) return PIKGPU_EINVAL;
  *s = b->sched;
  return PIKGPU_OK;
}

const char* pikgpu_batch_phase_name(int p, int* can_persist) {
  if (p < 0 || p >= PIKGPU_BATCH_PHASES) return nullptr;
  if (can_persist) *can_persist = kBatchPhases[p].loop ? 1 : 0;
  return kBatchPhases[p].name;
}

int pikgpu_batch_run(pikgpu_batch* b) {
  if (b == nullptr) return PIKGPU_EINVAL;
  if (b->n == 0) return BFail(b, PIKGPU_ESTATE, "no frames uploaded");
  BHIP_TRY(b, hipSetDevice(b->device));
  int rc;
  if (b->exec == nullptr && (rc = BuildGraph(b)) != PIKGPU_OK) return rc;
  BHIP_TRY(b, hipEventRecord(b->t0, b->main));
  BHIP_TRY(b, hipGraphLaunch(b->exec, b->main));
  BHIP_TRY(b, hipEventRecord(b->t1, b->main));
  b->ran = true;
  return PIKGPU_OK;
}

int pikgpu_batch_sync(pikgpu_batch* b) {
  if (b == nullptr) return PIKGPU_EINVAL;
  BHIP_TRY(b, hipSetDevice(b->device));
  BHIP_TRY(b, hipStreamSynchronize(b->main));
  if (b->ran) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->t0, b->t1) == hipSuccess) b->last_ms = ms;
  }
  return PIKGPU_OK;
}

float pikgpu_batch_last_ms(const pikgpu_batch* b) {
  return b ? b->last_ms : 0.0f;
}

int pikgpu_batch_download(pikgpu_batch* b, int i, float* out_planes[3],
                          size_t out_stride_bytes) {
  if (b == nullptr) return PIKGPU_EINVAL;
  if (!b->ran) return BFail(b, PIKGPU_ESTATE, "pikgpu_batch_run has not run");
  if (i < 0 || i >= b->n) return BFail(b, PIKGPU_EINVAL, "frame index");
  int rc = pikgpu_batch_sync(b);
  if (rc != PIKGPU_OK) return rc;
  rc = pikgpu_download(b->items[i], out_planes, out_stride_bytes);
  if (rc != PIKGPU_OK) return BFail(b, rc, b->items[i]->error);
  return PIKGPU_OK;
}

int pikgpu_reconstruct_batch(pikgpu_batch* b, const pikgpu_frame* frames,
                             int n, float* const* out_planes,
                             const size_t* out_stride_bytes) {
  int rc = pikgpu_batch_upload(b, frames, n);
  if (rc != PIKGPU_OK) return rc;
  if ((rc = pikgpu_batch_run(b)) != PIKGPU_OK) return rc;
  if (out_planes == nullptr || out_stride_bytes == nullptr) {
    return BFail(b, PIKGPU_EINVAL, "null outputs");
  }
  for (int i = 0; i < n; ++i) {
    float* planes[3] = {out_planes[3 * i], out_planes[3 * i + 1],
                        out_planes[3 * i + 2]};
    if ((rc = pikgpu_batch_download(b, i, planes, out_stride_bytes[i])) !=
        PIKGPU_OK) {
      return rc;
    }
  }
  return PIKGPU_OK;
}
