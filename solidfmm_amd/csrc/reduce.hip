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
constexpr int kBatch = 8;  // scratch loads in flight per lane

// One wave per target.  The items' descriptors are fetched 64 at a time (one
// per lane) and handed to the whole wave with readlane, and each lane keeps
// kBatch scratch loads in flight before it adds them, in list order, to its
// running sum: the latency of the item chain is hidden and the additions
// happen in exactly the order of the item list.
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
            for (uint32_t k0 = 0; k0 < T.count; k0 += 64) {
                const uint32_t nk = T.count - k0 < 64u ? T.count - k0 : 64u;
                // lane j holds the descriptor of item k0 + j
                uint64_t my_off = 0;
                uint32_t my_len = 0;
                if (uint32_t(lane) < nk) {
                    const RedItem it = items[T.first + k0 + uint32_t(lane)];
                    my_off = it.off;
                    my_len = it.len;
                }
                const uint32_t off_lo = uint32_t(my_off), off_hi = uint32_t(my_off >> 32);
                for (uint32_t k = 0; k < nk; k += kBatch) {
                    double v[kBatch];
                    bool on[kBatch];
#pragma unroll
                    for (int b = 0; b < kBatch; ++b) {
                        const uint32_t kk = k + uint32_t(b) < nk ? k + uint32_t(b) : nk - 1;
                        const uint64_t off = (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(off_hi), int(kk)))) << 32) |
                                             uint32_t(__builtin_amdgcn_readlane(int(off_lo), int(kk)));
                        const uint32_t len = uint32_t(__builtin_amdgcn_readlane(int(my_len), int(kk)));
                        on[b] = k + uint32_t(b) < nk && c < len;
                        v[b] = on[b] ? scratch[off + c] : 0.0;
                    }
#pragma unroll
                    for (int b = 0; b < kBatch; ++b)
                        if (on[b]) sum += v[b];
                }
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
