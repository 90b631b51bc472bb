/*
 * pikgpu.h -- C ABI of the MI355X (gfx950) PIK decoder-side reconstruction.
 *
 * The boundary sits where the reference decoder has finished entropy
 * decoding: the caller hands over the frame-wide quantized coefficients and
 * side information (exactly what the reference's host decoder holds in
 * FrameDecCache / GroupDecCache after DecodeDCGroups and DecodeFromBitstream)
 * and receives the linear-RGB Image3F that the reference's
 * FinalizeFrameDecoding leaves in `*idct`.
 *
 * Reference interfaces this replaces (google/pik):
 *   DequantImageAC          pik/compressed_image.h:83-87, .cc:827-867
 *   ReconOpsinImage         pik/compressed_image.h:90-97, .cc:936-1106
 *   FinalizeFrameDecoding   pik/compressed_image.h:102-110, .cc:1144-1181
 *   DequantDC + AdaptiveDCReconstruction
 *                           pik/compressed_dc.cc:181-230,
 *                           pik/adaptive_reconstruction.cc:300-332
 * and the per-group driver loop around them in PikPassToPixels
 * (pik/pik_frame.cc:1106-1149) and RoundtripImage
 * (pik/adaptive_quantization.cc:580-671).
 *
 * Conventions
 *  - Every function returns PIKGPU_OK (0) or a negative PIKGPU_E* code and
 *    never aborts on bad input; pikgpu_last_error() describes the failure.
 *  - A context owns one HIP stream and the device buffers of one frame; it is
 *    not internally locked (one context per host thread).
 *  - All image planes are row-major. Strides are in ELEMENTS unless named
 *    *_bytes.
 *  - Blocks are 8x8 pixels; xsize_blocks = ceil(xsize/8) etc. Colour tiles
 *    (CfL map, dequant control field) are 64x64 pixels.
 */
#ifndef PIKGPU_H_
#define PIKGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIKGPU_ABI_VERSION 3

#if defined(__GNUC__) || defined(__clang__)
#define PIKGPU_API __attribute__((visibility("default")))
#else
#define PIKGPU_API
#endif

enum {
  PIKGPU_OK = 0,
  PIKGPU_EINVAL = -1,    /* bad argument / inconsistent frame */
  PIKGPU_EHIP = -2,      /* HIP runtime error */
  PIKGPU_ENOMEM = -3,    /* device allocation failed */
  PIKGPU_ESTATE = -4,    /* call out of order (e.g. run before upload) */
  PIKGPU_EUNSUPPORTED = -5,
  PIKGPU_EDATA = -6      /* corrupt bitstream (host entropy decoding) */
};

/* AC strategy raw types; the per-block strategy byte is (type << 4) | index
 * of the 8x8 block inside a multi-block transform (pik/ac_strategy.h:55-75,
 * 303-313). */
enum {
  PIKGPU_ACS_DCT = 0,
  PIKGPU_ACS_IDENTITY = 1,
  PIKGPU_ACS_DCT2X2 = 2,
  PIKGPU_ACS_DCT4X4 = 3,
  PIKGPU_ACS_DCT16X16 = 4,
  PIKGPU_ACS_DCT32X32 = 5,
  PIKGPU_ACS_LINES = 6,
  PIKGPU_ACS_DCT_NOHF = 7,
  PIKGPU_ACS_DCT4X4_NOHF = 8
};

/* Gaborish strength, same numbering as GaborishStrength
 * (pik/pik_params.h:39-46). */
enum {
  PIKGPU_GABORISH_OFF = 0,
  PIKGPU_GABORISH_500 = 1,
  PIKGPU_GABORISH_750 = 2,
  PIKGPU_GABORISH_875 = 3,
  PIKGPU_GABORISH_1000 = 4
};

/* One dot-dictionary placement (BlockPosition + QuantizedBlock,
 * pik/block_dictionary.h:26-70). `pixels` holds the int8 patch for the three
 * channels, xsize*ysize used entries each, row-major. */
typedef struct pikgpu_dot {
  uint32_t x, y;          /* top-left corner in pixels */
  uint32_t xsize, ysize;  /* patch size, 1..8 */
  int32_t transform;      /* 0: copied as-is; 1: scaled/rotated splat */
  int32_t dx, dy, width;  /* only if transform: offset of top-right corner,
                             width in half-pixels */
  int8_t pixels[3][64];
} pikgpu_dot;

/* Frame-wide decoder state after entropy decoding. Host pointers. */
typedef struct pikgpu_frame {
  uint32_t xsize, ysize; /* image size in pixels (before 8-padding) */

  /* FrameHeader flags (pik/headers.h:182-294). */
  int32_t gaborish;            /* PIKGPU_GABORISH_* */
  int32_t predict_lf;          /* bool */
  int32_t predict_hf;          /* bool */
  int32_t adaptive_reconstruction; /* bool: EPF + AR clamp */
  int32_t epf_enable_adaptive; /* bool; only adaptive sigma is supported */
  int32_t epf_use_sharpened;   /* bool */

  /* Quantizer (pik/quantizer.h:111-389): serialized scalars. */
  int32_t global_scale; /* 1..32768 */
  int32_t quant_dc;     /* 1..256 */

  /* Quantized AC: 3 planes of ysize_blocks rows x (xsize_blocks*64) int16,
   * coefficients in the reference's frame-wide Image3S layout
   * (pik/compressed_image.cc:827-837). Coefficient 0 of each block is
   * ignored (DC arrives separately). */
  const int16_t* ac[3];
  size_t ac_stride;

  /* Quantized DC: 3 planes of ysize_blocks x xsize_blocks int16 (after DC
   * prediction has been undone by the host, i.e. GroupDecCache::quantized_dc
   * of pik/compressed_dc.cc:276-345). */
  const int16_t* dc[3];
  size_t dc_stride;

  /* Per-block side information, ysize_blocks x xsize_blocks each. */
  const int32_t* quant_field; size_t quant_field_stride; /* 1..256 */
  const uint8_t* ac_strategy; size_t ac_strategy_stride; /* (type<<4)|idx */
  const uint8_t* sigma_lut_ids; size_t sigma_lut_stride; /* 0 or 1 */

  /* Per 64x64 colour tile: ceil(ysize/64) x ceil(xsize/64). */
  const uint8_t* dequant_cf; size_t dequant_cf_stride;
  const int32_t* ytox_map; size_t ytox_stride; /* CfL factors, 0..255 */
  const int32_t* ytob_map; size_t ytob_stride;
  int32_t ytox_dc, ytob_dc;

  /* dequant_map[cf][qf-1] -> dequant table index (16 x 256 bytes). */
  const uint8_t* dequant_map;

  /* DequantMatrices (pik/quant_weights.h:202-266): the float table and the
   * offsets table indexed [(table * 7 + kind) * 3 + c]. */
  const float* dequant_table; size_t dequant_table_size;
  const uint32_t* dequant_offsets; size_t num_dequant_offsets;

  /* Dot dictionary placements, applied in order. */
  const pikgpu_dot* dots; size_t num_dots;

  /* FrameHeader::kGrayscaleOpt (pik/headers.h): the X and B DC of every
   * group are rebuilt from Y (GrayXyb::RestoreXB, pik/compressed_image.cc:
   * 94-209, 953-956) before the LF predictions, and the linear output is
   * set to its luma in all three channels (pik/pik_frame.cc:1155-1166).
   * bool. */
  int32_t grayscale;

  /* FrameHeader::flags verbatim (pik/headers.h:185-196, read at :209).
   * PIKGPU_FLAG_GRAYSCALE_OPT has the meaning of `grayscale` above (either
   * one selects it). PIKGPU_FLAG_GRADIENT_MAP (the DC is predicted from a
   * gradient map instead of AdaptiveDCReconstruction, pik/compressed_dc.cc:
   * 515-522) and PIKGPU_FLAG_NOISE (AddNoise after the final Gaborish,
   * pik/compressed_image.cc:1171-1174) are refused with
   * PIKGPU_EUNSUPPORTED, as is any bit the format does not define, so such
   * a frame never decodes silently without them. */
  uint32_t flags;
} pikgpu_frame;

/* Bits of pikgpu_frame.flags (FrameHeader::Flags, pik/headers.h:185-196). */
enum {
  PIKGPU_FLAG_GRADIENT_MAP = 1,
  PIKGPU_FLAG_GRAYSCALE_OPT = 2,
  PIKGPU_FLAG_NOISE = 4
};

typedef struct pikgpu_config {
  uint32_t abi_version; /* must be PIKGPU_ABI_VERSION */
  uint32_t flags;       /* reserved, 0 */
} pikgpu_config;

typedef struct pikgpu_ctx pikgpu_ctx;

/* Intermediate images that pikgpu_stage_dump can return after pikgpu_run.
 * Sizes: DC stages are xsize_blocks x ysize_blocks; COEFFS is the
 * (xsize_blocks*64) x ysize_blocks coefficient image after dequantisation
 * and predictions; pixel stages up to DOTS are 8-padded
 * (xsize_blocks*8 x ysize_blocks*8); LINEAR is xsize x ysize. */
enum {
  PIKGPU_STAGE_DC_DEQUANT = 0, /* DequantDC output */
  PIKGPU_STAGE_DC = 1,         /* after AdaptiveDCReconstruction */
  PIKGPU_STAGE_COEFFS = 2,     /* dequantised AC + LLF/LF/HF predictions */
  PIKGPU_STAGE_IDCT = 3,       /* ReconOpsinImage output */
  PIKGPU_STAGE_GUIDE = 4,      /* image the EPF guide bytes come from:
                                  ConvolveGaborish(idct), or idct with
                                  gaborish off or epf_use_sharpened */
  PIKGPU_STAGE_EPF = 5,        /* AdaptiveFilter output before AR clamp */
  PIKGPU_STAGE_AR = 6,         /* AdaptiveReconstruction output */
  PIKGPU_STAGE_DOTS = 7,       /* after BlockDictionary::AddTo */
  PIKGPU_STAGE_GABORISH = 8,   /* final ConvolveGaborish */
  PIKGPU_STAGE_LINEAR = 9,     /* OpsinToLinear output = final Image3F */
  PIKGPU_NUM_STAGES = 10
};

/* The library dequantisation tables, i.e. what DequantMatrices::Compute
 * (pik/quant_weights.cc:607-926) produces when the bitstream signals the
 * predefined tables, in the layout of pikgpu_frame's dequant_table /
 * dequant_offsets (one table set: 21 offsets, 4800 floats). Host-only.
 * Writes at most table_cap floats and offsets_cap offsets (either pointer
 * may be NULL to query the sizes) and returns PIKGPU_OK. */
PIKGPU_API int pikgpu_default_dequant_tables(float* table, size_t table_cap,
                                             uint32_t* offsets,
                                             size_t offsets_cap,
                                             size_t* table_size,
                                             size_t* num_offsets);

/* Replaces the ownership of FrameDecCache/GroupDecCache: creates a context on
 * HIP device `device`. Returns NULL on failure. */
PIKGPU_API pikgpu_ctx* pikgpu_create(int device, const pikgpu_config* config);
PIKGPU_API void pikgpu_destroy(pikgpu_ctx* ctx);

/* Message of the last failure on this context (never NULL). */
PIKGPU_API const char* pikgpu_last_error(const pikgpu_ctx* ctx);

/* Checks a frame without touching any device: sizes, strides, null arrays,
 * quantizer scalars, quant field range, AC strategy consistency (complete,
 * aligned multi-block transforms, as AcStrategyImage::SetFromRaw establishes,
 * pik/ac_strategy.cc:1066-1099), dequant table/offset bounds and dot
 * placements. The reference reports these as decoding failures (a false
 * Status from DecodeFromBitstream / DecodeDCGroups,
 * pik/compressed_image.cc:720-825). Returns PIKGPU_OK or the code
 * pikgpu_frame_upload would return; `msg` (may be NULL) receives the reason,
 * truncated to msg_size. */
PIKGPU_API int pikgpu_validate_frame(const pikgpu_frame* frame, char* msg,
                                     size_t msg_size);

/* Validates the frame and copies it to device memory (host -> HBM), growing
 * the context's buffers if needed. Replaces the caller-side filling of
 * FrameDecCache + frame-wide Image3S that precedes DequantImageAC. */
PIKGPU_API int pikgpu_frame_upload(pikgpu_ctx* ctx, const pikgpu_frame* frame);

/* Runs the whole reconstruction on the uploaded frame, device-resident,
 * asynchronously on the context stream: DequantDC, DC EPF, dequant + CfL,
 * predictions, variable IDCT, Gaborish guide, EPF, AR clamp, dots, Gaborish,
 * OpsinToLinear. Equivalent to the decode-side sequence DecodeDCGroups ->
 * per group {InitializeDecCache, DequantImageAC, ReconOpsinImage} ->
 * FinalizeFrameDecoding. `keep_stages` != 0 retains intermediates for
 * pikgpu_stage_dump (costs extra HBM traffic). */
PIKGPU_API int pikgpu_run(pikgpu_ctx* ctx, int keep_stages);

/* Outputs pikgpu_run produces (bit mask, default PIKGPU_OUTPUT_LINEAR). */
enum {
  PIKGPU_OUTPUT_LINEAR = 1, /* linear Image3F, 3 float planes in [0, 255] */
  PIKGPU_OUTPUT_SRGB8 = 2   /* 8-bit sRGB, interleaved R, G, B */
};

/* Selects the outputs of the following pikgpu_run calls. The 8-bit output
 * replaces the reference's conversion of the decoded image for an 8-bit
 * sRGB file: the colour transform of the linear image to the sRGB transfer
 * curve (ColorSpaceTransform, pik/color_management.cc, through lcms2), then
 * clamping to [0, 1], scaling to 255 and rounding (CastClip01 and
 * ToExternal, pik/external_image.cc:190-197, 666-696); it is computed in the
 * final kernel, so a run that asks only for it writes 3 bytes per pixel
 * instead of 12. */
PIKGPU_API int pikgpu_set_output(pikgpu_ctx* ctx, int outputs);

/* Adds an AC pass to the uploaded frame (multi-pass accumulation): after
 * pikgpu_frame_upload, call once per further pass, in pass order. ac[c] is
 * plane c of the pass's quantized coefficients in the layout of
 * pikgpu_frame.ac (rows of ac_stride int16). Replaces the decoding of
 * passes 1 .. num_passes - 1 with Dequant<false>::DoAC
 * (DecodeFromBitstream<false>, pik/pik_frame.cc:903-917;
 * pik/compressed_image.cc:575-703): each coefficient is the first pass's
 * dequantised value plus, in order, every later pass's own (the colour
 * correlation applied per pass). At most 16 passes; a new upload starts
 * again from one. pikgpu_fwd_adj takes single-pass frames only. */
PIKGPU_API int pikgpu_frame_add_pass(pikgpu_ctx* ctx, const int16_t* const ac[3],
                                     size_t ac_stride);

/* Encoder-side quantization of an XYB image under the uploaded frame's
 * side information (quant field, AC strategy, colour correlation, dequant
 * tables / control field / map, global_scale, quant_dc): the coefficient
 * half of RoundtripImage (pik/adaptive_quantization.cc:580-651), i.e.
 * InitializeFrameEncCache (pik/compressed_image.cc:210-304) and, per group,
 * ComputeCoefficients (:351-455). opsin[c]: plane c of the XYB image padded
 * to whole blocks (bx * 8 columns, by * 8 rows; rows of opsin_stride
 * floats; host memory). The quantized AC and DC replace the uploaded
 * frame's (later passes are dropped), so a following pikgpu_run decodes the
 * round trip; ac_out / dc_out (host, may be NULL) also receive them in the
 * layouts of pikgpu_frame.ac / .dc (rows of ac_stride / dc_stride int16).
 * With the frame's predict_lf / predict_hf flags the coefficients carry the
 * LF / HF prediction residuals as the reference encoder forms them
 * (PredictLfForEncoder, ComputePredictionResiduals, pik/ac_predictions.cc:
 * 477-557, 637-650). Frames with dots or the grayscale flag return
 * PIKGPU_EUNSUPPORTED, as does a quant value the encoder's dequant-map
 * lookup (indexed by the quant value itself, :417, 445) would read past the
 * map with. */
PIKGPU_API int pikgpu_quantize(pikgpu_ctx* ctx, const float* const opsin[3],
                               size_t opsin_stride, int16_t* ac_out[3],
                               size_t ac_stride, int16_t* dc_out[3],
                               size_t dc_stride);

/* Selects the downsampling factor of the following pikgpu_run calls: 1
 * (full size, the default), 2, 4 or 8. Replaces the `downsample` argument
 * of ReconOpsinImage and FinalizeFrameDecoding (pik/compressed_image.h:
 * 95-110; InverseIntegralTransform, pik/compressed_image.cc:869-934, and
 * FinalizeFrameDecoding, :1144-1181): each transform's pixels are averaged
 * over factor x factor squares, without adaptive reconstruction or Gaborish,
 * and the outputs of pikgpu_download / pikgpu_download_srgb8 are
 * ceil(xsize / factor) x ceil(ysize / factor). As in the reference a frame
 * with dots cannot be downsampled (pikgpu_run returns PIKGPU_EUNSUPPORTED),
 * and stage dumps need factor 1 (PIKGPU_EINVAL). */
PIKGPU_API int pikgpu_set_downsample(pikgpu_ctx* ctx, int factor);

/* Copies the 8-bit sRGB output of the last run (xsize x ysize pixels, rows
 * of 3 * xsize bytes) to host memory with the given row stride. */
PIKGPU_API int pikgpu_download_srgb8(pikgpu_ctx* ctx, uint8_t* rgb,
                                     size_t stride_bytes);

/* Waits for the context stream. */
PIKGPU_API int pikgpu_sync(pikgpu_ctx* ctx);

/* Copies the final linear Image3F (xsize x ysize, 3 planes) to host memory
 * with the given row stride in bytes. */
PIKGPU_API int pikgpu_download(pikgpu_ctx* ctx, float* out_planes[3],
                    size_t out_stride_bytes);

/* Convenience: upload + run + download (the drop-in for the decode-side
 * sequence listed above; the output matches what FinalizeFrameDecoding
 * leaves in *idct). */
PIKGPU_API int pikgpu_reconstruct(pikgpu_ctx* ctx, const pikgpu_frame* frame,
                       float* out_planes[3], size_t out_stride_bytes);

/* Size (width, height in elements) of a stage image for the uploaded frame. */
PIKGPU_API int pikgpu_stage_size(pikgpu_ctx* ctx, int stage, uint32_t* xsize,
                      uint32_t* ysize);

/* Copies stage `stage` (PIKGPU_STAGE_*) of the last pikgpu_run(ctx, 1) to
 * host planes of the size given by pikgpu_stage_size. */
PIKGPU_API int pikgpu_stage_dump(pikgpu_ctx* ctx, int stage, float* planes[3],
                      size_t stride_bytes);

/* Device-side timing of the last pikgpu_run (and, in the adj_* entries, of
 * the last pikgpu_fwd_adj), per kernel family, in milliseconds, measured
 * with HIP events on the context stream.
 * names[i] points to static strings. Returns the number of entries written
 * (<= max_entries). */
PIKGPU_API int pikgpu_last_timings(pikgpu_ctx* ctx, const char** names, float* ms,
                        int max_entries);

/* The HIP stream the context launches on (hipStream_t as void*). */
PIKGPU_API void* pikgpu_stream(pikgpu_ctx* ctx);

/* ------------------------------------------------------------------------
 * Batches: many frames of uneven sizes reconstructed by one launch. The
 * reference decodes one frame at a time (PikPassToPixels per frame,
 * pik/pik_frame.cc:1106-1181); a batch replaces a loop of those calls. Each
 * frame keeps its own device buffers; pikgpu_batch_run launches one hipGraph
 * holding, per phase of the reconstruction, one launch over the work of
 * every frame (see pikgpu_batch_schedule). Frames run at full size, linear
 * output, with the passes pikgpu_batch_add_pass gives them. */
typedef struct pikgpu_batch pikgpu_batch;

PIKGPU_API pikgpu_batch* pikgpu_batch_create(int device,
                                             const pikgpu_config* config);
PIKGPU_API void pikgpu_batch_destroy(pikgpu_batch* batch);
PIKGPU_API const char* pikgpu_batch_last_error(const pikgpu_batch* batch);
/* Validates and uploads n frames (host -> HBM); frame i's failure is
 * reported as "frame i: <reason>". */
PIKGPU_API int pikgpu_batch_upload(pikgpu_batch* batch,
                                   const pikgpu_frame* frames, int n);
/* Adds an AC pass to frame i of the uploaded batch (multi-pass
 * accumulation as pikgpu_frame_add_pass: Dequant<false>::DoAC of
 * DecodeFromBitstream<false>, pik/pik_frame.cc:903-917,
 * pik/compressed_image.cc:575-703); call once per further pass, in pass
 * order, after pikgpu_batch_upload. The next run rebuilds the graph. */
PIKGPU_API int pikgpu_batch_add_pass(pikgpu_batch* batch, int i,
                                     const int16_t* const ac[3],
                                     size_t ac_stride);
/* Launches the reconstruction of every uploaded frame (asynchronous). The
 * graph is built on the first run after an upload or added pass. */
PIKGPU_API int pikgpu_batch_run(pikgpu_batch* batch);
PIKGPU_API int pikgpu_batch_sync(pikgpu_batch* batch);
/* Device time of the last run in milliseconds (valid after a sync). */
PIKGPU_API float pikgpu_batch_last_ms(const pikgpu_batch* batch);
/* Linear Image3F of frame i of the last run (waits for the run). */
PIKGPU_API int pikgpu_batch_download(pikgpu_batch* batch, int i,
                                     float* out_planes[3],
                                     size_t out_stride_bytes);
/* How a batch runs (pikgpu_batch_set_schedule). PIKGPU_BATCH_QUEUE: one
 * launch per phase of the reconstruction over every frame of the batch,
 * each frame's work items (its plain launch's workgroups) placed one after
 * the other in the launch in `order`; a phase with persist[p] = k > 0 runs
 * k workgroups per CU, each looping over the work items (the phases whose
 * registers allow it; the others run one workgroup per item).
 * PIKGPU_BATCH_BRANCHES: every frame's own launch sequence, the frames
 * spread over parallel branches of the graph. The schedule is chosen by
 * measurement (pik_amd/autotune.py); the default is the queue with the
 * largest frames first and one workgroup per work item. */
#define PIKGPU_BATCH_PHASES 19
enum { PIKGPU_BATCH_QUEUE = 0, PIKGPU_BATCH_BRANCHES = 1 };
typedef struct pikgpu_batch_schedule {
  int32_t mode;   /* PIKGPU_BATCH_QUEUE or PIKGPU_BATCH_BRANCHES */
  int32_t order;  /* 0: upload order; 1: largest frame first; 2: smallest
                     first */
  int32_t persist[PIKGPU_BATCH_PHASES]; /* per phase, in the order of
                     pikgpu_batch_phase_name; 0 .. 16 */
} pikgpu_batch_schedule;
PIKGPU_API int pikgpu_batch_set_schedule(pikgpu_batch* batch,
                                         const pikgpu_batch_schedule* s);
PIKGPU_API int pikgpu_batch_get_schedule(const pikgpu_batch* batch,
                                         pikgpu_batch_schedule* s);
/* Name of phase p and whether it can run persistent (returns NULL past the
 * last phase). */
PIKGPU_API const char* pikgpu_batch_phase_name(int p, int* can_persist);

/* Upload + run + download: out_planes holds 3 plane pointers per frame,
 * out_stride_bytes one row stride per frame. */
PIKGPU_API int pikgpu_reconstruct_batch(pikgpu_batch* batch,
                                        const pikgpu_frame* frames, int n,
                                        float* const* out_planes,
                                        const size_t* out_stride_bytes);

/* ------------------------------------------------------------------------
 * Forward model and adjoint for a quant-field search. The reference's
 * encoder searches the quant field by repeated round trips, each a full
 * quantise + decode + Butteraugli (FindBestQuantization,
 * pik/adaptive_quantization.cc:713-852, decoding through RoundtripImage,
 * :580-671); it has no gradient. This evaluates the decode of the uploaded
 * frame as a function of a real-valued quant field qf (one value per 8x8
 * block) and the gradient of a loss with respect to qf:
 *   d      = AdjustQuantBias(q) * dequant matrix * (inv_global_scale / qf)
 *            (the table stays the one the frame's integer field selects)
 *   linear = OpsinToLinear(crop(Gaborish(IDCT(LLF/LF + CfL(d)))))
 * i.e. the frame's decode with HF prediction, adaptive reconstruction and
 * dots switched off; with qf equal to the integer field, `linear` is bit for
 * bit pikgpu_run's output for a frame with those three flags off.
 *
 * The loss is sum (linear - target)^2 when target[0] is set (dL/dlinear =
 * 2 (linear - target)); otherwise dl_dlinear supplies dL/dlinear. */
enum { PIKGPU_ADJ_DEVICE_POINTERS = 1 };

typedef struct pikgpu_adj_io {
  uint32_t flags;               /* PIKGPU_ADJ_DEVICE_POINTERS: every pointer
                                   below is device memory and the call is
                                   asynchronous on pikgpu_stream(ctx) */
  const float* qf;              /* ysize_blocks rows, finite and > 0 */
  size_t qf_stride;             /* floats per row */
  const float* target[3];       /* xsize x ysize linear planes, or NULL */
  const float* dl_dlinear[3];   /* used when target[0] is NULL */
  size_t in_stride;             /* floats per row of target / dl_dlinear */
  float* linear[3];             /* optional forward output (all or none) */
  size_t linear_stride;         /* floats per row */
  float* dl_dqf;                /* ysize_blocks rows of xsize_blocks */
  size_t dqf_stride;            /* floats per row */
  float* dl_dcoeffs[3];         /* optional dL/dd per channel, before CfL
                                   (zero at the LLF positions): ysize_blocks
                                   rows of xsize_blocks * 64 (all or none) */
  size_t dcoeffs_stride;        /* floats per row */
  double* loss;                 /* optional, with target: the loss */
} pikgpu_adj_io;

/* One forward + adjoint evaluation on the uploaded frame. Host pointers:
 * returns when the outputs are written. Device pointers: enqueued on the
 * context stream (pikgpu_sync waits); qf values are not checked. */
PIKGPU_API int pikgpu_fwd_adj(pikgpu_ctx* ctx, const pikgpu_adj_io* io);

/* ------------------------------------------------------------------------
 * Gradient quant-field search. Replaces the reference encoder's
 * Butteraugli-driven search (FindBestQuantization, pik/adaptive_quantization.
 * cc:677-853: twelve round trips, each a quantization, a decode and a
 * Butteraugli comparison, rescaling the field by tile distances) by Adam
 * steps on the log quant field driven by the forward + adjoint of the decode
 * (pikgpu_fwd_adj's model), all device-resident on the context stream.
 *
 * Each step quantizes `opsin` with the current integer field (as
 * pikgpu_quantize), evaluates the squared linear-RGB error against `target`
 * and its gradient dL/dd with respect to the dequantized coefficients, and
 * attributes the error to the blocks to first order: D_b = 1/2 sum_k g_k e_k
 * (e: each coefficient's quantization error). With the noise model
 * D ~ qf^-2 the gradient of the rate-distortion objective with respect to a
 * transform's log field is sum over its blocks of (lambda - 2 w_b D_b);
 * Adam follows it, the field being rounded to integers in [1, qf_max] for
 * every quantization. The stationary point equalizes w_b D_b at lambda / 2.
 *
 * The uploaded frame supplies everything but the field's values: strategy,
 * colour correlation, dequant tables / control field / map, global_scale,
 * quant_dc, header flags (frames without LF / HF prediction, dots or the
 * grayscale flag, as pikgpu_quantize) and the initial integer field. On
 * return the context holds the frame quantized with the final field, so
 * pikgpu_run decodes it. */
typedef struct pikgpu_search_params {
  int32_t steps;         /* Adam steps (the field is quantized steps + 1
                            times) */
  float learning_rate;   /* Adam step size on log qf */
  float beta1, beta2, eps;
  float lambda;          /* rate weight per block; <= 0 selects the mean of
                            2 w_b D_b over the blocks at the initial field
                            times -lambda (0: that mean itself), i.e. the
                            distortion level of the start is kept */
  int32_t qf_max;        /* largest integer quant value, <= 256 */
  float trust_low;       /* each transform's field stays within
                            [trust_low, trust_high] times its initial value
                            (0 < trust_low <= 1 <= trust_high), as the
                            reference search bounds its field by the
                            initial one's range */
  float trust_high;
} pikgpu_search_params;

typedef struct pikgpu_search_io {
  const float* opsin[3];   /* XYB image as pikgpu_quantize takes it (padded
                              to whole blocks; host) */
  size_t opsin_stride;     /* floats per row */
  const float* target[3];  /* linear RGB original, xsize x ysize (host) */
  size_t target_stride;    /* floats per row */
  const float* weight;     /* per-block perceptual weights w_b (ysize_blocks
                              rows of xsize_blocks; host), or NULL for 1 */
  size_t weight_stride;
  int32_t* quant_field;    /* out (host, may be NULL): final integer field */
  size_t quant_field_stride;
  int16_t* ac[3];          /* out (host, all or none): final quantized AC, */
  size_t ac_stride;        /* layout of pikgpu_frame.ac */
  int16_t* dc[3];          /* out (host, all or none): quantized DC */
  size_t dc_stride;
  double* loss;            /* out (host, steps + 1 entries, may be NULL):
                              squared linear-RGB error at each quantization */
  double* rate;            /* out (host, steps + 1, may be NULL): sum of
                              log(qf) over the blocks at each quantization */
  float* lambda_used;      /* out (may be NULL): the rate weight applied */
} pikgpu_search_io;

PIKGPU_API int pikgpu_search_quant_field(pikgpu_ctx* ctx,
                                         const pikgpu_search_params* params,
                                         const pikgpu_search_io* io);

/* ------------------------------------------------------------------------
 * Reconstruction without device allocation. No counterpart in the
 * reference (its decoder allocates its images per frame); this entry point
 * is for callers that cannot allocate device memory: code running while a
 * HIP graph is captured around it, tight serving loops, harnesses that
 * forbid hipMalloc. It does what pikgpu_reconstruct does (upload + run +
 * download of the linear Image3F, same bits) without a context: every
 * device buffer of the frame lives in one fixed array of
 * PIKGPU_NOALLOC_SCRATCH_BYTES per device, which the library holds from its
 * first use on a device until the process ends, and the call makes no
 * hipMalloc / hipFree. A frame that needs more scratch than that (about
 * 2.7 megapixels; 1920x1080 fits) is refused with PIKGPU_EINVAL before the
 * device is touched.
 *
 * Runs on the current device, on `stream` (a hipStream_t; NULL: the null
 * stream), and returns when the output is in out_planes (host memory, rows
 * of out_stride_bytes). Calls on one device share its array, which keeps
 * the last call's contents: calls that overlap on one device race, so the
 * caller serialises them. `msg` (may be NULL) receives the reason of a
 * failure, truncated to msg_size. */
#define PIKGPU_NOALLOC_SCRATCH_BYTES (128u << 20)
PIKGPU_API int pikgpu_reconstruct_noalloc(const pikgpu_frame* frame,
                                          float* out_planes[3],
                                          size_t out_stride_bytes,
                                          void* stream, char* msg,
                                          size_t msg_size);

/* ------------------------------------------------------------------------
 * Host entropy decoding of one pass's AC groups into the frame-wide int16
 * AC planes that pikgpu_frame.ac (and pikgpu_frame_add_pass) take.
 *
 * Replaces the entropy half of DecodeFromBitstream<first>
 * (pik/compressed_image.cc:719-813): DecodeNoise, DecodeCoeffOrder x 3,
 * DecodeHistograms, then DecodeAC per 64x64 tile with GatherCoefficients
 * for 16x16 and 32x32 transforms (pik/noise.cc:529-539,
 * pik/entropy_coder.cc:822-836, 889-924, 1149-1212, pik/ac_strategy.h:
 * 254-272). The dequantization the reference does per tile runs on the GPU
 * for the whole frame instead. Runs on the host only (no device calls).
 *
 * Groups are 512x512 pixels, row-major; group i's bytes start at its
 * noise parameters (the start of its pass data after the group header,
 * pik/pik_frame.cc:902-915). The AC strategy bytes are those of
 * pikgpu_frame.ac_strategy. Every block of every group is written. Groups
 * decode in parallel on `threads` host threads (<= 0: one per hardware
 * thread). Returns PIKGPU_OK, PIKGPU_EINVAL for inconsistent arguments, or
 * PIKGPU_EDATA when a group's bitstream is invalid (the lowest such group
 * is named in `msg`); the AC of a failed group is unspecified. */
typedef struct pikgpu_ac_pass {
  uint32_t xsize, ysize;          /* frame size in pixels */
  const uint8_t* ac_strategy;     /* ceil(ysize/8) rows of ceil(xsize/8) */
  size_t ac_strategy_stride;
  size_t num_groups;              /* ceil(xsize/512) * ceil(ysize/512) */
  const uint8_t* const* group_data;
  const size_t* group_size;       /* bytes of each group's pass data */
  int16_t* ac[3];                 /* out: X, Y, B; rows of >= 64 * bx */
  size_t ac_stride;
  float* noise;                   /* out (may be NULL): alpha, gamma, beta
                                     per group (NoiseParams) */
  size_t* group_consumed;         /* out (may be NULL): bytes read per
                                     group */
} pikgpu_ac_pass;

PIKGPU_API int pikgpu_decode_ac_pass(const pikgpu_ac_pass* pass, int threads,
                                     char* msg, size_t msg_size);

/* ------------------------------------------------------------------------
 * Host decoding of a frame's DC groups section into the quantized DC and
 * the per-block side information that pikgpu_frame takes (dc, ac_strategy,
 * quant_field, sigma_lut_ids).
 *
 * Replaces DecodeDCGroups (pik/compressed_dc.cc:436-528) up to DequantDC:
 * the size table of the 256x256-block DC groups (SizeCoderT<0x150F0E0C>),
 * then per group DecodeDCGroup (:276-345) -- the DC residual image
 * (DecodeImage, pik/entropy_coder.cc:839-939), the adaptive DC predictor
 * undone (ExpandDC, :156-197, pik/dc_predictor.cc), and under one ANS state
 * DecodeAcStrategy, DecodeQuantField and DecodeARParameters
 * (pik/entropy_coder.cc:944-1147). DequantDC and the DC smoothing run on
 * the GPU (pikgpu_run). Host only; groups decode on `threads` host threads
 * (<= 0: one per hardware thread). The old DC coder only (the default;
 * use_new_dc is refused upstream), no multi-pass hints, no gradient map.
 * `data` starts at the size table (byte aligned, after the colour maps);
 * `consumed` receives the section's length in bytes. Returns PIKGPU_OK,
 * PIKGPU_EINVAL or PIKGPU_EDATA (the lowest failing group in `msg`). */
typedef struct pikgpu_dc_section {
  uint32_t xsize, ysize;            /* frame size in pixels */
  const uint8_t* data;
  size_t size;
  int16_t* dc[3];                   /* out: quantized DC, X Y B */
  size_t dc_stride;
  uint8_t* ac_strategy;             /* out: (type << 4) | block index */
  size_t ac_strategy_stride;
  int32_t* quant_field;             /* out */
  size_t quant_field_stride;
  uint8_t* sigma_lut_ids;           /* out */
  size_t sigma_lut_stride;
  size_t consumed;                  /* out */
} pikgpu_dc_section;

PIKGPU_API int pikgpu_decode_dc_groups(pikgpu_dc_section* section,
                                       int threads, char* msg,
                                       size_t msg_size);

/* ------------------------------------------------------------------------
 * Host parsing of one whole frame, from its FrameHeader to the end of its
 * last pass, into the pikgpu_frame the reconstruction takes.
 *
 * Replaces the bitstream half of PikPassToPixels (pik/pik_frame.cc:
 * 935-1187) and PikGroupToPixels (:846-931) for the default decoder
 * parameters (full size, every pass): ReadPassHeader (FrameHeader,
 * pik/headers.h:182-294, through the field coders of pik/fields.h:67-300,
 * 853-935), DequantMatrices::Decode (pik/quant_weights.cc:590-603),
 * Quantizer::Decode (pik/quantizer.cc:63-71), DecodeColorMap for Y-to-B and
 * Y-to-X (pik/color_correlation.cc:244-262), the DC groups section
 * (pikgpu_decode_dc_groups), BlockDictionary::Decode (pik/block_dictionary.
 * cc:199-307), DecodeDequantControlField / DecodeDequantControlFieldMap
 * (pik/quant_weights.cc:942-958, 1054-1075), then per pass the group size
 * table and per group its GroupHeader (first pass only, pik/headers.h:
 * 106-136) and AC data (pikgpu_decode_ac_pass).
 *
 * `data` starts at the byte-aligned FrameHeader and runs to the end of the
 * file (the reference's reader spans the whole file); xsize / ysize are the
 * FileHeader's image size (the file header itself, with its colour
 * metadata, is not parsed here). Groups decode on `threads` host threads
 * (<= 0: one per hardware thread). *out is NULL or a parsed frame of an
 * earlier call, whose storage is then reused (its contents are unspecified
 * after a failure, when *out is left as it was); on success *out receives
 * the parsed frame, which owns every array; release it with
 * pikgpu_parsed_free.
 *
 * Returns PIKGPU_OK; PIKGPU_EINVAL for bad arguments; PIKGPU_EDATA for an
 * invalid frame, wherever the reference decoder fails and where it would
 * index outside its own tables (a dequant control field value >= 16, a
 * dequant map entry naming a missing table, a dot outside the image, a
 * dot transform of zero size, length or width, no passes); or
 * PIKGPU_EUNSUPPORTED for frames outside this decoder: lossless frames,
 * alpha, the gradient map, custom (non-library) quantization tables, more
 * than 16 passes, or a dot dictionary of more than 2^20 patches or 2^22
 * placements (the reference would allocate gigabytes). `msg` names the
 * failing section. The colour maps' own failures are ignored, as the
 * reference ignores DecodeColorMap's result (pik/pik_frame.cc:1049-1050):
 * a map whose code does not parse keeps its default. */
typedef struct pikgpu_parsed pikgpu_parsed;

typedef struct pikgpu_parsed_info {
  uint64_t frame_size;        /* FrameHeader::size (bytes; 0 if unset) */
  uint32_t num_passes;
  uint32_t num_groups;        /* 512x512 groups per pass */
  uint32_t epf_sigma;         /* EpfParams::sigma (if not adaptive) */
  uint32_t num_downsampling;  /* downsampling_factor_to_passes entries */
  const uint32_t* downsampling; /* pairs {factor, passes} */
  const float* noise;         /* num_passes x num_groups x {alpha, gamma,
                                 beta} (NoiseParams per group and pass) */
  size_t consumed;            /* bytes from the FrameHeader to the end of
                                 the last pass's last group */
} pikgpu_parsed_info;

PIKGPU_API int pikgpu_parse_frame(const uint8_t* data, size_t size,
                                  uint32_t xsize, uint32_t ysize, int threads,
                                  pikgpu_parsed** out, char* msg,
                                  size_t msg_size);
/* Fills `frame` (pointers into the parsed frame's storage, valid until it
 * is freed; ac is the first pass) and `info` (either may be NULL). */
PIKGPU_API int pikgpu_parsed_get(const pikgpu_parsed* parsed,
                                 pikgpu_frame* frame, pikgpu_parsed_info* info);
/* The quantized AC of pass `pass` (0 .. num_passes - 1) in the layout of
 * pikgpu_frame.ac: ac[c] receives plane c, *ac_stride the row stride. */
PIKGPU_API int pikgpu_parsed_pass_ac(const pikgpu_parsed* parsed,
                                     uint32_t pass, const int16_t* ac[3],
                                     size_t* ac_stride);
PIKGPU_API void pikgpu_parsed_free(pikgpu_parsed* parsed);

/* The whole decode of one frame on a context: pikgpu_parse_frame, then
 * pikgpu_frame_upload, pikgpu_frame_add_pass for every further pass and
 * pikgpu_run -- PikPassToPixels end to end. The result is read with
 * pikgpu_download / pikgpu_download_srgb8; *consumed (may be NULL) receives
 * the frame's length in bytes. Returns the first failing step's code; the
 * message is pikgpu_last_error(ctx). */
PIKGPU_API int pikgpu_decode_frame(pikgpu_ctx* ctx, const uint8_t* data,
                                   size_t size, uint32_t xsize,
                                   uint32_t ysize, int threads,
                                   size_t* consumed);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* PIKGPU_H_ */
