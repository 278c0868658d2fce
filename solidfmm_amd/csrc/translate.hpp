// Internal launcher interface of the translation kernels (translate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/solidfmm_gpu.h"

namespace sfmm {

// True when this build has a kernel for (op, p_in, p_out) in `family`.
bool have_kernel(int op, uint32_t p_in, uint32_t p_out, int family = 0);

// Highest order with a kernel.
size_t max_kernel_order();

// Workgroups for a persistent kernel: as many as can be resident on the
// current device at once, but no more than `work` (one per unit of work).
uint32_t persistent_grid(const void *kernel, int block, size_t work);

// Kernel families: one lane per translation (FP64 VALU, orders up to 12)
// and one wave per 16 translations (swaps on the FP64 matrix cores).
// and the generic kernel (one workgroup per translation, any orders).
enum Family { kAuto = 0, kLaneKernels = 1, kMfmaKernels = 2, kGenericKernel = 3 };

namespace mfma {
inline constexpr int kMfmaMin = 1;
inline constexpr int kMfmaMax = 24;
hipError_t launch_m2m(uint32_t P, const sfmm_item *, size_t, const double *, double *, const double *, hipStream_t);
hipError_t launch_m2l(uint32_t P, const sfmm_item *, size_t, const double *, double *, const double *, hipStream_t);
hipError_t launch_l2l(uint32_t P, const sfmm_item *, size_t, const double *, double *, const double *, hipStream_t);
}  // namespace mfma

// The generic kernel: items of any orders up to `pc`.
hipError_t launch_generic(int op, uint32_t pc, const sfmm_item *items, size_t n, const double *src, double *dst,
                          hipStream_t stream);

// Enqueue dst += T(src) for n device-resident items, all of orders
// (p_in, p_out) and with pairwise distinct targets.
// `family` picks the kernel family (kAuto: the lane kernels where they
// exist, else the MFMA kernels).
hipError_t launch_translate(int op, uint32_t p_in, uint32_t p_out, const sfmm_item *items, size_t n,
                            const double *src, double *dst, hipStream_t stream, int family = kAuto);

}  // namespace sfmm
