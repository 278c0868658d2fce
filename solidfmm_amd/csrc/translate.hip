// Dispatch of the translation kernels: (op, kernel choice) -> launch.
//
// The kernels themselves live in lane_kernel.hpp (one lane per
// translation), translate_mfma.hpp (16 translations per wave on the FP64
// matrix cores), generic.hip (one workgroup per translation, any order) and,
// for the reduced-precision modes, lane_prec.hpp, mfma_prec.hpp (FP32
// matrix cores) and prec_generic.hip.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/solidfmm_gpu.h"
#include "lane_kernel.hpp"
#include "mfma_tables.hpp"
#include "translate.hpp"

namespace sfmm {

size_t max_kernel_order() { return size_t(mfma::kMaxOrder); }

bool family_covers(int family, uint32_t pb)
{
    if (pb < 1) return false;
    switch (family) {
    case kLaneKernels: return pb <= kLaneMax;
    case kMfmaKernels: return pb >= uint32_t(mfma::kMfmaMin) && pb <= uint32_t(mfma::kMfmaMax);
    case kGenericKernel: return pb <= uint32_t(mfma::kMaxOrder);
    default: return false;
    }
}

int default_family(uint32_t pb)
{
    if (pb <= kLaneAutoMax) return kLaneKernels;
    if (pb <= uint32_t(mfma::kMfmaMax)) return kMfmaKernels;
    return kGenericKernel;
}

bool have_kernel(int op, uint32_t p_in, uint32_t p_out, int family)
{
    if (op < 0 || op > 2 || p_in < 1 || p_out < 1) return false;
    const uint32_t pb = p_in > p_out ? p_in : p_out;
    if (family == kAuto) return pb <= uint32_t(mfma::kMaxOrder);
    return family_covers(family, pb);
}

uint32_t persistent_grid(const void *kernel, int block, size_t work)
{
    struct Key {
        const void *k;
        int dev;
    };
    static std::mutex mu;
    static std::vector<std::pair<Key, uint32_t>> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    uint32_t resident = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (auto &e : cache)
            if (e.first.k == kernel && e.first.dev == dev) resident = e.second;
        if (!resident) {
            int per_cu = 0, cus = 0;
            // a failed query falls back to one workgroup per CU; its error
            // is cleared so the launch that follows does not report it
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) {
                (void)hipGetLastError();
                per_cu = 1;
            }
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
                (void)hipGetLastError();
                cus = 1;
            }
            resident = uint32_t(per_cu) * uint32_t(cus);
            cache.push_back({{kernel, dev}, resident});
        }
    }
    if (work < resident) return uint32_t(work < 1 ? 1 : work);
    return resident;
}

// Items per wave of a lane launch of n items: 64, unless the last round of
// resident waves would be partly empty -- then the items are spread evenly
// over whole rounds (1e5 items on 1,024 resident waves: 2,041 waves of 49
// instead of 1,563 of 64, two full rounds either way).
uint32_t lane_items_per_wave(const void *kernel, size_t n)
{
    const size_t slots = persistent_grid(kernel, 64, size_t(1) << 40);
    const size_t waves = (n + 63) / 64;
    if (slots == 0 || waves <= slots) return 64;
    const size_t rounds = (waves + slots - 1) / slots;
    const size_t ipw = (n + rounds * slots - 1) / (rounds * slots);
    return uint32_t(ipw < 1 ? 1 : ipw > 64 ? 64 : ipw);
}

namespace {

template <bool RT>
hipError_t launch_mfma(int op, uint32_t pk, const sfmm_item *items, size_t n, const double *src, double *dst,
                       const double *table, hipStream_t stream, int acc)
{
    switch (op) {
    case SFMM_M2M: return mfma::launch_op<SFMM_M2M, RT>(pk, items, n, src, dst, table, stream, acc);
    case SFMM_M2L: return mfma::launch_op<SFMM_M2L, RT>(pk, items, n, src, dst, table, stream, acc);
    default: return mfma::launch_op<SFMM_L2L, RT>(pk, items, n, src, dst, table, stream, acc);
    }
}

}  // namespace

hipError_t launch_kernel(int op, const KernelChoice &k, const sfmm_item *items, size_t n, const double *src,
                         double *dst, hipStream_t stream)
{
    const int acc = k.acc ? 1 : 0;
    if (k.precision != SFMM_FP64) {
        if (k.family == kLaneKernels && op == SFMM_M2L && !k.rt)
            return launch_prec_lane_m2l(k.precision, k.pk, items, n, src, dst, stream, acc);
        if (k.family == kMfmaKernels && op == SFMM_M2L && !k.rt && k.precision == SFMM_FP32)
            return launch_prec_mfma_m2l(k.pk, items, n, src, dst, stream, acc);
        if (k.family == kGenericKernel)
            return launch_prec_generic(k.precision, op, k.pk, items, n, src, dst, stream, acc);
        return hipErrorInvalidValue;
    }
    if (op < 0 || op > 2 || !family_covers(k.family, k.pk)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    switch (k.family) {
    case kLaneKernels: {
        // scratch launches with runs use the per-item-order kernels (the
        // fixed-order ones have no code for runs)
        const bool runs = !k.acc && k.seg;
        const lane::Table &t = (k.rt || runs) ? lane::rt_table() : lane::fixed_table();
        lane::kernel_fn f = t.f[op][k.pk - 1];
        const uint32_t block = 64;
        uint32_t ipw = lane_items_per_wave(reinterpret_cast<const void *>(f), n);
        if (runs) ipw = 64;  // the plan lays runs out in 64-item waves
        const uint32_t grid = static_cast<uint32_t>((n + ipw - 1) / ipw);
        hipLaunchKernelGGL(f, dim3(grid), dim3(block), 0, stream, src, dst, items, static_cast<uint32_t>(n),
                           runs ? 2 : acc, ipw);
        return hipGetLastError();
    }
    case kMfmaKernels: {
        const double *table = mfma::device_table();
        if (!table) return hipErrorOutOfMemory;
        const int mode = (!k.acc && k.seg) ? 2 : acc;
        return k.rt ? launch_mfma<true>(op, k.pk, items, n, src, dst, table, stream, mode)
                    : launch_mfma<false>(op, k.pk, items, n, src, dst, table, stream, mode);
    }
    default: return launch_generic(op, k.pk, items, n, src, dst, stream, acc);
    }
}

hipError_t launch_translate(int op, uint32_t p_in, uint32_t p_out, const sfmm_item *items, size_t n,
                            const double *src, double *dst, hipStream_t stream, int family)
{
    if (!have_kernel(op, p_in, p_out, family)) return hipErrorInvalidValue;
    const uint32_t pb = p_in > p_out ? p_in : p_out;
    const int fam = family == kAuto ? default_family(pb) : family;
    return launch_kernel(op, KernelChoice{fam, pb, p_in != p_out}, items, n, src, dst, stream);
}

}  // namespace sfmm
