// Internal launcher interface of the translation kernels (translate.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/solidfmm_gpu.h"

namespace sfmm {

// True when this build has a kernel for (op, p_in, p_out).
bool have_kernel(int op, uint32_t p_in, uint32_t p_out);

// Highest order with a kernel.
size_t max_kernel_order();

// Enqueue dst += T(src) for n device-resident items, all of orders
// (p_in, p_out) and with pairwise distinct targets.
hipError_t launch_translate(int op, uint32_t p_in, uint32_t p_out, const sfmm_item *items, size_t n,
                            const double *src, double *dst, hipStream_t stream);

}  // namespace sfmm
