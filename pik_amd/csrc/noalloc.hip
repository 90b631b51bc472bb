// The no-allocation entry point (pikgpu_reconstruct_noalloc, pikgpu.h):
// the reconstruction of pikgpu_reconstruct for callers that must not
// allocate device memory. It has no counterpart in the reference, whose
// decoder allocates its images per frame.
//
// Every device buffer of the frame (its uploaded inputs, the intermediates
// and the output staging) is carved out of `scratch`, one file-scope array
// per device, at offsets computed on the host; the constant tables sit at
// its start. The array's name is a host-side stand-in: kernels and copies
// use the address hipGetSymbolAddress returns for the current device (taken
// once per device, on the first call there), never the name. Calls on one
// device share the array, which keeps the last call's contents: calls that
// overlap on one device race.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "host_api.h"
#include "kernels.h"

using namespace pikgpu;

namespace {

size_t DivCeil(size_t a, size_t b) { return (a + b - 1) / b; }

// Zero fill of n floats.
__global__ void __launch_bounds__(256) KFillZero(float* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

constexpr size_t kNoAllocBytes = size_t(PIKGPU_NOALLOC_SCRATCH_BYTES);
constexpr size_t kNoAllocAlign = 256;
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

struct NoAllocDevice {
  unsigned char* ptr = nullptr;  // device address of `scratch`
  bool ready = false;
};
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

size_t AlignUp(size_t v) { return (v + kNoAllocAlign - 1) & ~(kNoAllocAlign - 1); }

// Offsets of one frame's buffers inside the scratch array: a bump
// allocator over [base, base + kNoAllocBytes); with base 0 it measures.
struct Carve {
  uintptr_t base;
  size_t off;
  template <class T>
  T* Take(size_t count) {
    const size_t at = off;
    off = AlignUp(off + (count ? count : 1) * sizeof(T));
    return reinterpret_cast<T*>(base + at);
  }
};

// Lays the frame's buffers out from `base` (the constants first) and
// returns the bytes they span. Host-side lists: the pred2x2 classes and the
// dot bins.
size_t NoAllocLayout(const pikgpu_frame* f, const FrameParams& P,
                     size_t pred_count, size_t bins, uintptr_t base,
                     FrameBuffers* B, DeviceConsts** K) {
  const size_t nblk = size_t(P.bx) * P.by;
  const size_t ngroups = size_t(P.gx) * P.gy;
  const size_t ipitch = DivCeil(size_t(P.px), 32) * 32;
  const size_t ntiles = size_t(P.tx) * P.ty;
  const size_t grows = DivCeil(size_t(P.py), 32) * 32 + 16;
  const size_t gpitch = DivCeil(size_t(P.px), 32) * 32 + 64;
  Carve c{base, 0};
  *K = c.Take<DeviceConsts>(1);
  memset(B, 0, sizeof(*B));
  B->ac = c.Take<int16_t>(3 * nblk * 64);
  B->dc16 = c.Take<int16_t>(3 * nblk);
  B->qf = c.Take<int32_t>(nblk);
  B->acs = c.Take<uint8_t>(nblk);
  B->lut = c.Take<uint8_t>(nblk);
  B->dcf = c.Take<uint8_t>(ntiles);
  B->ytox = c.Take<int32_t>(ntiles);
  B->ytob = c.Take<int32_t>(ntiles);
  B->dqmap = c.Take<uint8_t>(16 * 256);
  B->table = c.Take<float>(f->dequant_table_size);
  B->offsets = c.Take<uint32_t>(f->num_dequant_offsets);
  B->dots = c.Take<pikgpu_dot>(f->num_dots);
  B->dot_tile_ids = c.Take<int32_t>(bins);
  B->pred_list = c.Take<uint32_t>(pred_count);
  B->desc = c.Take<uint4>(2 * nblk);
  B->dcq = c.Take<float>(3 * nblk);
  B->dc = c.Take<float>(3 * nblk);
  B->llf = c.Take<float>(ngroups * 3 * kLlfPlane);
  B->tmp2x2 = c.Take<float>(ngroups * 3 * kG2Plane);
  B->lf2x2 = c.Take<float>(ngroups * 3 * kG2Plane);
  B->pred2x2 = c.Take<float>(ngroups * 3 * kG2Plane);
  B->idct = c.Take<float>(3 * ipitch * P.py);
  B->ar = c.Take<float>(3 * ipitch * P.py);
  B->gimg = c.Take<uint8_t>(3 * grows * gpitch);
  B->mm_partial = c.Take<float>(size_t(kMinMaxMaxGroups) * 6);
  B->mm = c.Take<float>(8);
  // KFinal reads `ar` only, so the linear output reuses the IDCT image
  // (3 planes of ysize x opitch <= 3 planes of py x ipitch)
  B->out = B->idct;
  B->gpitch = int(gpitch);
  B->grows = int(grows);
  B->ipitch = int(ipitch);
  B->opitch = int(DivCeil(size_t(f->xsize), 32) * 32);
  return c.off;
}

// Set-up of the current device, once: the device address of `scratch` and
// the constant tables at its start.
int NoAllocDeviceInit(int dev, unsigned char** ptr, std::string* err) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoAllocDevice& d = g_noalloc[dev];
  if (!d.ready) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      *err = std::string("hipGetSymbolAddress(scratch): ") +
             hipGetErrorString(e);
      return PIKGPU_EHIP;
    }
    DeviceConsts h;
    memset(&h, 0, sizeof(h));
    HostTables host;
    FillScaleTables(&h.tscale[0][0][0], &h.ascale[0][0][0]);
    ComputeHostTables(&host);
    memcpy(h.mul_table, host.mul_table, sizeof(h.mul_table));
    OpsinInverse(h.opsin_inv);
    AdjointNorms(h.adj_norm);
    ComputeGrayXyb(&h.gray);
    e = hipMemcpy(p, &h, sizeof(h), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      *err = std::string("constant tables: ") + hipGetErrorString(e);
      return PIKGPU_EHIP;
    }
    d.ptr = static_cast<unsigned char*>(p);
    d.ready = true;
  }
  *ptr = d.ptr;
  return PIKGPU_OK;
}

int NoAllocRun(const pikgpu_frame* f, float* const out_planes[3],
               size_t out_stride_bytes, hipStream_t st, std::string* err) {
  int rc = ValidateFrame(f, err);
  if (rc != PIKGPU_OK) return rc;
  if (out_planes == nullptr || !out_planes[0] || !out_planes[1] ||
      !out_planes[2]) {
    *err = "null output plane";
    return PIKGPU_EINVAL;
  }
  if (out_stride_bytes < size_t(f->xsize) * 4) {
    *err = "output stride smaller than a row";
    return PIKGPU_EINVAL;
  }
  FrameParams P;
  FillParams(f, &P);
  std::vector<uint32_t> pred_list;
  PredClasses(f, &pred_list, P.pred_end);
  std::vector<int32_t> dot_ids, dot_starts, dot_list;
  if (f->num_dots) {
    BinDots(f, P.px, P.py, &dot_ids, &dot_starts, &dot_list);
  }
  const size_t bins = dot_ids.size() + dot_starts.size() + dot_list.size();
  FrameBuffers B;
  DeviceConsts* K = nullptr;
  const size_t need =
      NoAllocLayout(f, P, pred_list.size(), bins, 0, &B, &K);
  if (need > kNoAllocBytes) {
    *err = "frame needs " + std::to_string(need) +
           " bytes of device scratch; the no-allocation entry point has " +
           std::to_string(kNoAllocBytes);
    return PIKGPU_EINVAL;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
    *err = "no usable current device";
    return PIKGPU_EHIP;
  }
  unsigned char* base = nullptr;
  if ((rc = NoAllocDeviceInit(dev, &base, err)) != PIKGPU_OK) return rc;
  NoAllocLayout(f, P, pred_list.size(), bins, reinterpret_cast<uintptr_t>(base),
                &B, &K);
  B.num_tables = int(f->num_dequant_offsets / (kNumQuantKinds * 3));
  P.dot_tiles = int(dot_ids.size());
  P.dot_tiles_x = int(DivCeil(size_t(P.px), 32));
  P.out_flags = kOutLinear;
  const size_t nb = size_t(P.bx) * P.by;
  const size_t bx = size_t(P.bx), by = size_t(P.by);
  const size_t tx = size_t(P.tx), ty = size_t(P.ty);
#define TRY(expr)                                                    \
  do {                                                               \
    if ((e = (expr)) != hipSuccess) {                                \
      *err = std::string(#expr) + ": " + hipGetErrorString(e);       \
      return PIKGPU_EHIP;                                            \
    }                                                                \
  } while (0)
  auto up2d = [&](const void* dst, const void* src, size_t spitch,
                  size_t row, size_t rows) {
    return Copy2D(const_cast<void*>(dst), row, src, spitch, row,
                            rows, hipMemcpyHostToDevice, st);
  };
  for (int c = 0; c < 3; ++c) {
    TRY(up2d(B.ac + c * nb * 64, f->ac[c], f->ac_stride * 2, bx * 128, by));
    TRY(up2d(B.dc16 + c * nb, f->dc[c], f->dc_stride * 2, bx * 2, by));
  }
  TRY(up2d(B.qf, f->quant_field, f->quant_field_stride * 4, bx * 4, by));
  TRY(up2d(B.acs, f->ac_strategy, f->ac_strategy_stride, bx, by));
  TRY(up2d(B.lut, f->sigma_lut_ids, f->sigma_lut_stride, bx, by));
  TRY(up2d(B.dcf, f->dequant_cf, f->dequant_cf_stride, tx, ty));
  TRY(up2d(B.ytox, f->ytox_map, f->ytox_stride * 4, tx * 4, ty));
  TRY(up2d(B.ytob, f->ytob_map, f->ytob_stride * 4, tx * 4, ty));
  auto up = [&](const void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(const_cast<void*>(dst), src, bytes,
                                  hipMemcpyHostToDevice, st)
                 : hipSuccess;
  };
  TRY(up(B.dqmap, f->dequant_map, 16 * 256));
  TRY(up(B.table, f->dequant_table, f->dequant_table_size * 4));
  TRY(up(B.offsets, f->dequant_offsets, f->num_dequant_offsets * 4));
  TRY(up(B.pred_list, pred_list.data(), pred_list.size() * 4));
  if (f->num_dots) {
    int32_t* bin = const_cast<int32_t*>(B.dot_tile_ids);
    TRY(up(B.dots, f->dots, f->num_dots * sizeof(pikgpu_dot)));
    TRY(up(bin, dot_ids.data(), dot_ids.size() * 4));
    TRY(up(bin + dot_ids.size(), dot_starts.data(), dot_starts.size() * 4));
    TRY(up(bin + dot_ids.size() + dot_starts.size(), dot_list.data(),
           dot_list.size() * 4));
    B.dot_tile_start = bin + dot_ids.size();
    B.dot_list = bin + dot_ids.size() + dot_starts.size();
  } else {
    B.dot_tile_ids = nullptr;
  }
  // the pred2x2 positions no block writes must read zero (ZeroFillImage,
  // pik/compressed_image.cc:976); the array keeps the last call's
  // contents, so they are cleared on every call, by a kernel
  {
    const size_t n = size_t(P.gx) * P.gy * 3 * kG2Plane;
    KFillZero<<<DivCeil(n, 256), 256, 0, st>>>(B.pred2x2, n);
  }
  TRY(EnqueueFullSize(P, B, K, st, nullptr));
  const size_t oplane = size_t(B.opitch) * P.ysize;
  for (int c = 0; c < 3; ++c) {
    TRY(Copy2D(out_planes[c], out_stride_bytes, B.out + c * oplane,
                         size_t(B.opitch) * 4, size_t(P.xsize) * 4, P.ysize,
                         hipMemcpyDeviceToHost, st));
  }
  TRY(hipStreamSynchronize(st));
#undef TRY
  return PIKGPU_OK;
}

}  // namespace

extern "C" int pikgpu_reconstruct_noalloc(const pikgpu_frame* frame,
                                          float* out_planes[3],
                                          size_t out_stride_bytes,
                                          void* stream, char* msg,
                                          size_t msg_size) {
  std::string err;
  const int rc = NoAllocRun(frame, out_planes, out_stride_bytes,
                            static_cast<hipStream_t>(stream), &err);
  if (msg && msg_size) {
    snprintf(msg, msg_size, "%s", rc == PIKGPU_OK ? "" : err.c_str());
  }
  return rc;
}
