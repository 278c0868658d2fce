// Internal launcher interface of the translation kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/solidfmm_gpu.h"

namespace sfmm {

// Kernel families:
//   lane     one lane per translation, FP64 VALU, orders up to 12
//            (lane_kernel.hpp);
//   MFMA     one wave per 16 translations, swaps on the FP64 matrix cores,
//            orders up to 24 (translate_mfma.hpp);
//   generic  one workgroup per translation, any orders up to 86
//            (generic.hip).
// Lane and MFMA kernels are instantiated per square order PB; their RT
// variants take items of any orders p_in, p_out <= PB.
enum Family { kAuto = 0, kLaneKernels = 1, kMfmaKernels = 2, kGenericKernel = 3 };

inline constexpr uint32_t kLaneMax = 12;
// lane kernels by default up to this order: at 11 they beat MFMA on 1e5
// independent translations (M2L 1.18e9 against 1.04e9/s, M2M 1.37e9
// against 1.15e9, L2L 1.36e9 against 1.05e9); at 12 MFMA wins (M2L 8.9e8
// against 6.1e8)
inline constexpr uint32_t kLaneAutoMax = 11;

namespace mfma {
inline constexpr int kMfmaMin = 1;
inline constexpr int kMfmaMax = 24;
template <int OP, bool RT>
hipError_t launch_op(uint32_t P, const sfmm_item *items, size_t n, const double *src, double *dst,
                     const double *table, hipStream_t stream, int acc);
}  // namespace mfma

// Highest order with a kernel.
size_t max_kernel_order();

// True when `family` (not kAuto) has a kernel for items whose larger order
// is pb.
bool family_covers(int family, uint32_t pb);

// The family kAuto picks for items whose larger order is pb.
int default_family(uint32_t pb);

// True when the library has a kernel for (op, p_in, p_out) in `family`.
bool have_kernel(int op, uint32_t p_in, uint32_t p_out, int family = kAuto);

// Workgroups for a persistent kernel: as many as can be resident on the
// current device at once, but no more than `work` (one per unit of work).
uint32_t persistent_grid(const void *kernel, int block, size_t work);

// Reduced-precision kernels (arithmetic policies of prec.hpp): a lane
// kernel for square M2L up to kPrecLaneMax and a generic kernel for every
// operator and order up to kPrecMaxOrder.
inline constexpr uint32_t kPrecLaneMax = 16;    // FP32
inline constexpr uint32_t kPrecLaneMaxDW = 10;  // double-word FP32
inline constexpr uint32_t kPrecMaxOrder = 24;
hipError_t launch_prec_lane_m2l(int precision, uint32_t P, const sfmm_item *items, size_t n, const double *src,
                                double *dst, hipStream_t stream, int acc);
hipError_t launch_prec_generic(int precision, int op, uint32_t pc, const sfmm_item *items, size_t n,
                               const double *src, double *dst, hipStream_t stream, int acc);
// FP32 square M2L on the matrix cores, orders 1..kPrecMaxOrder
// (mfma_prec.hip).  FP32 contexts use it above kPrecLaneAutoMax.
inline constexpr uint32_t kPrecLaneAutoMax = 16;
hipError_t launch_prec_mfma_m2l(uint32_t P, const sfmm_item *items, size_t n, const double *src, double *dst,
                                hipStream_t stream, int acc);

// One launch of a plan: family (not kAuto), the kernel's order bound pk,
// whether items carry their own orders (rt), the arithmetic (SFMM_FP64,
// SFMM_FP32, SFMM_SPLIT32) and whether targets accumulate (acc) or are
// overwritten with the translation (scratch blocks of a segmented
// reduction).  The generic kernels always read per-item orders; pk is then
// the largest order of the group.
struct KernelChoice {
    int family;
    uint32_t pk;
    bool rt;
    int precision = SFMM_FP64;
    bool acc = true;
    // scratch mode with runs of one target's items sharing a block (FP64
    // lane kernels only; the plan lays the runs out)
    bool seg = false;
};

// Segmented reduction: target block t += sum of its items' scratch blocks,
// in list order.  Both arrays are device memory.
struct RedTarget {
    uint64_t dst;    // offset (reals) of the target block in the target pool
    uint32_t first;  // first entry of the target's items in RedItem[]
    uint32_t count;  // number of items
    uint32_t len;    // reals to add (the longest of the items' blocks)
    uint32_t pad;
};
struct RedItem {
    uint64_t off;  // offset (reals) of the item's block in the scratch pool
    uint32_t len;  // reals the translation wrote
    uint32_t pad;
};
hipError_t launch_reduce(const RedTarget *targets, size_t n_targets, const RedItem *items, const double *scratch,
                         double *dst, hipStream_t stream);

// Enqueue dst += T(src) for n device-resident items with pairwise distinct
// targets and orders within `k`.
hipError_t launch_kernel(int op, const KernelChoice &k, const sfmm_item *items, size_t n, const double *src,
                         double *dst, hipStream_t stream);

// The generic kernel: items of any orders up to `pc`.
hipError_t launch_generic(int op, uint32_t pc, const sfmm_item *items, size_t n, const double *src, double *dst,
                          hipStream_t stream, int acc = 1);

// Particle-side operators over device-resident cells (particles.hip).
// dim: values per point and blocks per cell, 1..kMaxParticleDim.
inline constexpr int kMaxParticleDim = 4;
hipError_t launch_p2m(const double *pts, const sfmm_cell *cells, size_t n_cells, double *pool, int dim,
                      hipStream_t stream);
hipError_t launch_l2p(const double *pts, const sfmm_cell *cells, size_t n_cells, const double *pool, double *phi,
                      double *grad, int dim, hipStream_t stream);

// Enqueue dst += T(src) for n items all of orders (p_in, p_out) with
// pairwise distinct targets, in `family` (kAuto: default_family).
hipError_t launch_translate(int op, uint32_t p_in, uint32_t p_out, const sfmm_item *items, size_t n,
                            const double *src, double *dst, hipStream_t stream, int family = kAuto);

}  // namespace sfmm
