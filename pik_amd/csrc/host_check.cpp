// Host build of the device transforms and tables (libpikgpu_host.so), so the
// CPU test suite can check them bit for bit against the reference without a
// GPU. Compiled with the same -ffp-contract=off as the device code.
#include <string.h>

#include "common.h"
#include "xform.h"

#define HAPI extern "C" __attribute__((visibility("default")))

// Transform of one block of raw strategy `raw`: dir 0 = inverse
// (coefficients -> pixels), 1 = forward. Both sides are contiguous
// N x N blocks (N = 8, 16, 32).
HAPI int pikgpu_host_block_transform(int raw, int dir, const float* in,
                                     float* out) {
  const int n = 8 * pikgpu::AcsCovered(raw);
  if (dir == 0) {
    pikgpu::BlockInverse(raw, in, out, n);
  } else {
    pikgpu::BlockForward(raw, in, n, out);
  }
  return 0;
}

// Transpose of the inverse transform of one block (contiguous N x N).
HAPI int pikgpu_host_block_adjoint(int raw, const float* in, float* out) {
  float norms[pikgpu::kAdjNormFloats];
  pikgpu::AdjointNorms(norms);
  const int n = 8 * pikgpu::AcsCovered(raw);
  pikgpu::BlockAdjoint(raw, norms, in, n, out);
  return 0;
}

HAPI void pikgpu_host_adjoint_norms(float* out) {
  pikgpu::AdjointNorms(out);
}

HAPI void pikgpu_host_mul_table(int* out) {
  pikgpu::HostTables t;
  pikgpu::ComputeHostTables(&t);
  memcpy(out, t.mul_table, sizeof(t.mul_table));
}

HAPI void pikgpu_host_scale_tables(float* tscale, float* ascale) {
  pikgpu::FillScaleTables(tscale, ascale);
}

HAPI void pikgpu_host_gaborish_weights(int strength, float* w) {
  pikgpu::GaborishWeights(strength, w);
}

HAPI void pikgpu_host_ub4(float* k) {
  float t[3][4];
  pikgpu::Ub4Kernel(t);
  memcpy(k, t, sizeof(t));
}

// GrayXyb's table: out[0] = sub, out[1] = mul, then 16 each of x_slope,
// x_const, b_slope, b_const (66 doubles).
HAPI void pikgpu_host_gray_xyb(double* out) {
  pikgpu::GrayXybTable t;
  pikgpu::ComputeGrayXyb(&t);
  out[0] = t.sub;
  out[1] = t.mul;
  for (int i = 0; i < pikgpu::kGrayPieces; ++i) {
    out[2 + i] = t.x_slope[i];
    out[18 + i] = t.x_const[i];
    out[34 + i] = t.b_slope[i];
    out[50 + i] = t.b_const[i];
  }
}

HAPI void pikgpu_host_linear_to_xyb(float r, float g, float b, float* xyb) {
  pikgpu::LinearToXybApprox(r, g, b, &xyb[0], &xyb[1], &xyb[2]);
}
