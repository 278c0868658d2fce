// Instantiations of the MFMA translation kernels for M2L (square orders
// kMfmaMin..kMfmaMax) and their launcher.
#include "translate_mfma.hpp"

namespace sfmm {
namespace mfma {

namespace {
using fn_t = void (*)(const double *, double *, const sfmm_item *, uint32_t, const double *);

template <int... Ps>
constexpr auto make_table(std::integer_sequence<int, Ps...>)
{
    struct T {
        fn_t f[sizeof...(Ps)];
    } t{{&k_translate_mfma<SFMM_M2L, Ps + kMfmaMin, Ps + kMfmaMin>...}};
    return t;
}
}  // namespace

hipError_t launch_m2l(uint32_t P, const sfmm_item *items, size_t n, const double *src, double *dst,
                         const double *table, hipStream_t stream)
{
    static const auto tab = make_table(std::make_integer_sequence<int, kMfmaMax - kMfmaMin + 1>{});
    if (P < uint32_t(kMfmaMin) || P > uint32_t(kMfmaMax)) return hipErrorInvalidValue;
    const fn_t f = tab.f[P - kMfmaMin];
    const size_t groups = (n + 15) / 16;
    const uint32_t grid = persistent_grid(reinterpret_cast<const void *>(f), 64 * kWaves, (groups + kWaves - 1) / kWaves);
    hipLaunchKernelGGL(f, dim3(grid), dim3(64 * kWaves), 0, stream, src, dst, items, static_cast<uint32_t>(n), table);
    return hipGetLastError();
}

}  // namespace mfma
}  // namespace sfmm
