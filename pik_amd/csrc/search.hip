// Gradient quant-field search (pikgpu_search_quant_field): the kernels of
// one Adam step on the log quant field, between the encoder's quantization
// (KEncQuant) and the forward + adjoint of the decode (KAdjForward /
// KAdjBackward).
//
// The reference searches the field by repeated round trips judged by
// Butteraugli (FindBestQuantization, pik/adaptive_quantization.cc:677-853);
// it has no gradient. Here every step quantizes the image with the current
// integer field, decodes it with the differentiable model and back-
// propagates the squared linear-RGB error to the dequantized coefficients
// (g = dL/dd). To first order the error the quantization of block b causes
// is D_b = 1/2 sum_k g_k e_k, with e = d - x the quantization error of each
// coefficient (x the encoder's unquantized value): for L = |J d - t|^2,
// g = 2 J^T J e, so the D_b sum to e^T J^T J e, and each block's share
// carries the decode's weighting (IDCT, colour correlation, Gaborish, the
// opsin cube) of its own coefficients. Under the quantization-noise model
// D ~ step^2 ~ qf^-2, dD_b / d log qf = -2 D_b; the rate grows with
// log qf. Each transform's log field u then follows Adam on
//   G = sum over its blocks of ( -2 w_b D_b + lambda ),
// whose stationary point equalizes the weighted distortion w_b D_b across
// blocks at the level lambda sets (w: per-block perceptual weights given by
// the caller, 1 by default).
#include "kernels.h"

namespace pikgpu {

// Integer field from the log field: each block takes its transform's value
// (stored at the transform's first block), rounded and clamped to
// [1, qmax]; `qfr` receives it as float for the adjoint. Per workgroup the
// sum of log(qf) (the rate proxy) goes to `part`.
__global__ void __launch_bounds__(256)
    KSearchRound(FrameParams P, FrameBuffers B, const float* u, int qmax,
                 int32_t* qf, float* qfr, float* part) {
  __shared__ float s[4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float lq = 0.0f;
  if (i < P.bx * P.by) {
    const int y = i / P.bx, x = i - y * P.bx;
    const uint8_t a = B.acs[i];
    const int sc = AcsCovered(a >> 4), idx = a & 15;
    const int fy = y - idx / sc, fx = x - idx % sc;
    const float v = rintf(__expf(u[fy * P.bx + fx]));
    const int q = int(fminf(fmaxf(v, 1.0f), float(qmax)));
    qf[i] = q;
    qfr[i] = float(q);
    lq = __logf(float(q));
  }
  for (int m = 32; m >= 1; m >>= 1) lq += __shfl_xor(lq, m, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = lq;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// Distortion attributed to each storage block (one wave per block, lane k
// = coefficient slot k; `coef` and `grad` hold three planes in the AC
// layout): D_b = 1/2 sum over k and the three channels of
// g e, in the total (after colour correlation) coefficient domain. The
// adjoint's gradients for X and B are already in that domain; Y's total
// gradient is its pre-correlation one minus the correlation factors times
// X's and B's. LLF slots have g = 0.
__global__ void __launch_bounds__(256)
    KSearchDist(FrameParams P, FrameBuffers B, const float* coef,
                const float* grad, float* dist) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = threadIdx.x & 63;
  if (b >= P.bx * P.by) return;
  const int y = b / P.bx, x = b - y * P.bx;
  const size_t plane = size_t(P.by) * P.bx * 64;
  const size_t at = size_t(b) * 64 + k;
  const BlockDq dx = MakeBlockDq(P, B, 0, y, x);
  const BlockDq dy = MakeBlockDq(P, B, 1, y, x);
  const BlockDq db = MakeBlockDq(P, B, 2, y, x);
  const float vy = DequantWith(dy, 1, B.ac[plane + at], k, 0.0f);
  const float vx = DequantWith(dx, 0, B.ac[at], k, vy);
  const float vb = DequantWith(db, 2, B.ac[2 * plane + at], k, vy);
  const float gx = grad[at], gyt = grad[plane + at], gb = grad[2 * plane + at];
  const float gy = gyt - dx.cc * gx - db.cc * gb;
  float d = gx * (vx - coef[at]) + gy * (vy - coef[plane + at]) +
            gb * (vb - coef[2 * plane + at]);
  for (int m = 32; m >= 1; m >>= 1) d += __shfl_xor(d, m, 64);
  if (k == 0) dist[b] = 0.5f * d;
}

// One Adam step (step number t >= 1) on each transform's log field, at its
// first block: gradient sum over the covered blocks of -2 w D + lambda;
// u clamped to [ulo, uhi] and to the trust region [u0 + dlo, u0 + dhi]
// around its initial value u0.
__global__ void __launch_bounds__(256)
    KSearchUpdate(FrameParams P, FrameBuffers B, SearchArgs S, int t,
                  const float* dist, const float* u0, float* u, float* m,
                  float* v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.bx * P.by) return;
  const uint8_t a = B.acs[i];
  if ((a & 15) != 0) return;
  const int sc = AcsCovered(a >> 4);
  const int y = i / P.bx, x = i - y * P.bx;
  float g = 0.0f;
  for (int iy = 0; iy < sc; ++iy) {
    for (int ix = 0; ix < sc; ++ix) {
      const int j = (y + iy) * P.bx + x + ix;
      const float w = S.weight ? S.weight[j] : 1.0f;
      g += *S.lambda - 2.0f * w * dist[j];
    }
  }
  const float mi = S.beta1 * m[i] + (1.0f - S.beta1) * g;
  const float vi = S.beta2 * v[i] + (1.0f - S.beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(S.beta1, float(t)));
  const float vh = vi / (1.0f - powf(S.beta2, float(t)));
  const float ui = u[i] - S.lr * mh / (sqrtf(vh) + S.eps);
  const float lo = fmaxf(S.ulo, u0[i] + S.dlo);
  const float hi = fminf(S.uhi, u0[i] + S.dhi);
  u[i] = fminf(fmaxf(ui, lo), hi);
}

// The rate weight: `fixed` when positive, else kappa times the mean over
// the blocks of 2 w_b D_b (one workgroup).
__global__ void __launch_bounds__(256)
    KSearchLambda(const float* dist, const float* weight, int n, float fixed,
                  float kappa, float* out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    acc += 2.0 * double(weight ? weight[i] : 1.0f) * double(dist[i]);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out = fixed > 0.0f ? fixed : float(kappa * s[0] / double(n > 0 ? n : 1));
  }
}

// Initial state: u = u0 = log of the frame's integer field, Adam moments
// zero.
__global__ void __launch_bounds__(256)
    KSearchInit(FrameParams P, FrameBuffers B, float* u0, float* u, float* m,
                float* v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.bx * P.by) return;
  u[i] = __logf(float(B.qf[i]));
  u0[i] = u[i];
  m[i] = 0.0f;
  v[i] = 0.0f;
}

}  // namespace pikgpu
