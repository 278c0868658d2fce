// FP32 instantiations of the reduced-precision lane M2L kernel (orders
// 1..kPrecLaneMax) and the launcher of both precision lane families.
#include "lane_prec.hpp"

namespace sfmm {
namespace lprec {

kernel_fn kernel_f32(uint32_t P)
{
    static const auto t = make_table<prec::F32>(std::make_integer_sequence<int, int(kPrecLaneMax)>{});
    return (P >= 1 && P <= kPrecLaneMax) ? t.f[P - 1] : nullptr;
}

}  // namespace lprec

hipError_t launch_prec_lane_m2l(int precision, uint32_t P, const sfmm_item *items, size_t n, const double *src,
                                double *dst, hipStream_t stream, int acc)
{
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    lprec::kernel_fn f = nullptr;
    switch (precision) {
    case SFMM_FP32: f = lprec::kernel_f32(P); break;
    case SFMM_SPLIT32: f = lprec::kernel_dw(P); break;
    default: break;
    }
    if (!f) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const uint32_t block = 64;
    const uint32_t grid = static_cast<uint32_t>((n + block - 1) / block);
    hipLaunchKernelGGL(f, dim3(grid), dim3(block), 0, stream, src, dst, items, static_cast<uint32_t>(n), acc);
    return hipGetLastError();
}

}  // namespace sfmm
