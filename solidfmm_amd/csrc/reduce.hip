// Segmented reduction for translation lists whose targets repeat many times
// (an FMM cell receives ~190 M2L contributions from its interaction list).
//
// The translation kernels write each item's result into its own scratch
// block (no read of the target); this kernel then adds, for every target,
// the scratch blocks of its items in list order and adds the sum into the
// target once.  One wave per target, lanes over the coefficients (64
// consecutive reals per pass, so every scratch read is coalesced); the
// order of the additions is fixed, so results are bitwise reproducible.
// The reference applies items sharing a target one after another
// (multithreaded_translator.cpp:298-319 sends all of a target's items to one
// thread); here the sum of a target's contributions is formed first and
// added to the target once, which changes only the rounding.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "translate.hpp"

namespace sfmm {
namespace {

constexpr int kWaves = 4;

__global__ void __launch_bounds__(64 * kWaves)
    k_reduce(const RedTarget *__restrict__ targets, uint32_t n_targets, const RedItem *__restrict__ items,
             const double *__restrict__ scratch, double *__restrict__ dst)
{
    const int lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t t = blockIdx.x * kWaves + wave; t < n_targets; t += gridDim.x * kWaves) {
        const RedTarget T = targets[t];
        for (uint32_t c0 = 0; c0 < T.len; c0 += 64) {
            const uint32_t c = c0 + uint32_t(lane);
            double sum = 0.0;
            for (uint32_t k = 0; k < T.count; ++k) {
                const RedItem it = items[T.first + k];
                if (c < it.len) sum += scratch[it.off + c];
            }
            if (c < T.len) dst[T.dst + c] += sum;
        }
    }
}

}  // namespace

hipError_t launch_reduce(const RedTarget *targets, size_t n_targets, const RedItem *items, const double *scratch,
                         double *dst, hipStream_t stream)
{
    if (n_targets == 0) return hipSuccess;
    if (n_targets > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t grid =
        persistent_grid(reinterpret_cast<const void *>(&k_reduce), 64 * kWaves, (n_targets + kWaves - 1) / kWaves);
    hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(64 * kWaves), 0, stream, targets, static_cast<uint32_t>(n_targets),
                       items, scratch, dst);
    return hipGetLastError();
}

}  // namespace sfmm
