// FP32 M2L on the matrix cores (mfma_prec.hpp): kernels of orders 1..16
// and the launcher.
#include "mfma_prec.hpp"

namespace sfmm {
namespace mprec {

fn_t kernel_lo(uint32_t P)
{
    static const auto fns = make_fns<1>(std::make_integer_sequence<int, 16>{});
    return (P >= 1 && P <= 16) ? fns.f[P - 1] : nullptr;
}

}  // namespace mprec

hipError_t launch_prec_mfma_m2l(uint32_t P, const sfmm_item *items, size_t n, const double *src, double *dst,
                                hipStream_t stream, int acc)
{
    using namespace mprec;
    if (P < 1 || P > uint32_t(kMax) || n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    const fn_t f = P <= 16 ? kernel_lo(P) : P <= 20 ? kernel_mid(P) : kernel_hi(P);
    if (!f) return hipErrorInvalidValue;
    const size_t groups = (n + 15) / 16;
    const uint32_t grid =
        persistent_grid(reinterpret_cast<const void *>(f), 64 * kWaves, (groups + kWaves - 1) / kWaves);
    hipLaunchKernelGGL(f, dim3(grid), dim3(64 * kWaves), 0, stream, src, dst, items, static_cast<uint32_t>(n), acc);
    return hipGetLastError();
}

}  // namespace sfmm
