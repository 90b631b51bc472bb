/*
 * pik_oracle.h -- CPU restatement of the PIK decode-side reconstruction.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark compare the HIP library
 * against; the product (pik_amd) never links or calls it.
 *
 * Single-threaded plain C, frame at a time, in the reference scalar build's
 * floating-point operation order (compile with -ffp-contract=off), so that
 * its output can be compared bit for bit with the reference built from its
 * own sources (oracle/_ref) and with the GPU.
 */
#ifndef PIK_ORACLE_H_
#define PIK_ORACLE_H_

#include "../include/pikgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reconstructs `frame` into the linear Image3F `out` (3 planes of
 * xsize x ysize floats, row-major, no padding). When `stages` is non-NULL,
 * stages[3 * s + c] (s = PIKGPU_STAGE_*) receives plane c of stage s in the
 * sizes pikgpu_stage_size reports, rows contiguous; NULL entries are
 * skipped. Returns 0, or a negative PIKGPU_E* code with a message in
 * pikoracle_last_error(). */
int pikoracle_recon(const pikgpu_frame* frame, float* const* stages,
                    float* const out[3]);

/* pikoracle_recon with AC passes after the frame's own (multi-pass
 * accumulation): extra_ac[3 * p + c] is plane c of extra pass p, rows of
 * extra_stride int16 in the frame's AC layout. */
int pikoracle_recon_passes(const pikgpu_frame* frame, int num_extra,
                           const int16_t* const* extra_ac, size_t extra_stride,
                           float* const* stages, float* const out[3]);

/* The downsampled decode for ds = 2, 4, 8 (no AR, no Gaborish, no dots):
 * `out` receives 3 planes of ceil(xsize / ds) x ceil(ysize / ds) floats. */
/* Encoder-side quantization of an 8-padded XYB image (three dense
 * bx*8 x by*8 planes) under the frame's side information, as
 * InitializeFrameEncCache + ComputeCoefficients compute it
 * (pik/compressed_image.cc:210-455): quantized AC (three dense planes of
 * by rows x bx*64) and DC (three dense bx x by planes). Frames without
 * predictions, dots or the grayscale flag. */
int pikoracle_quantize(const pikgpu_frame* frame, const float* const* opsin,
                       int16_t* const* ac_out, int16_t* const* dc_out);

int pikoracle_recon_downsampled(const pikgpu_frame* frame, int ds,
                                float* const out[3]);

const char* pikoracle_last_error(void);

/* One block transform of raw strategy `raw`: dir 0 = TransformToPixels,
 * 1 = TransformFromPixels; contiguous N x N in and out (N = 8, 16, 32). */
int pikoracle_block_transform(int raw, int dir, const float* in, float* out);

#ifdef __cplusplus
}
#endif

#endif /* PIK_ORACLE_H_ */
