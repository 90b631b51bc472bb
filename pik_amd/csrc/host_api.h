// Host-side pieces of the C ABI shared by its translation units (api.hip,
// noalloc.hip): frame validation, derived frame parameters, host-built
// lists and the full-size launch sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device.h"

namespace pikgpu {

// Most workgroups of a KMinMaxPartial launch (6 floats of partials each).
constexpr int kMinMaxMaxGroups = 4096;

// Checks that a frame is complete and self-consistent (what the reference
// rejects as a decoding error); host-only. PIKGPU_OK or the error code,
// with the reason in *err.
int ValidateFrame(const pikgpu_frame* f, std::string* err);

// Geometry and derived scalars of a validated frame.
void FillParams(const pikgpu_frame* f, FrameParams* P);

// First blocks grouped by the class of their pred2x2 computation; end[k]
// is the end of class k in `list`.
void PredClasses(const pikgpu_frame* f, std::vector<uint32_t>* list,
                 int end[4]);

// Dots binned into 32x32 tiles of the padded image (px x py): tile ids, CSR
// starts and dot indices in application order.
void BinDots(const pikgpu_frame* f, int px, int py, std::vector<int32_t>* ids,
             std::vector<int32_t>* starts, std::vector<int32_t>* list);

// Inverse of the opsin absorbance matrix (Inv3x3Matrix<float>).
void OpsinInverse(float out[9]);

// A 2-D copy, issued as one linear copy when both sides' rows are
// contiguous. The runtime moves a host-side 2-D copy row by row (from
// pageable memory, one staged copy per row: a 4096x4096 frame's 512-row
// side planes took ~30 ms that way, against ~2 ms for the whole upload as
// linear copies).
hipError_t Copy2D(void* dst, size_t dpitch, const void* src, size_t spitch,
                  size_t row_b, size_t rows, hipMemcpyKind kind,
                  hipStream_t stream);

// The full-size reconstruction of a frame whose inputs are in device memory
// (B), asynchronous on `st`; allocates nothing. `ev` (may be null) receives
// events 0..7 between the kernel families.
hipError_t EnqueueFullSize(const FrameParams& P, const FrameBuffers& B,
                           const DeviceConsts* K, hipStream_t st,
                           const hipEvent_t* ev);

}  // namespace pikgpu
