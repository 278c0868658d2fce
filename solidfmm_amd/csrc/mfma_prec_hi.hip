// FP32 M2L on the matrix cores (mfma_prec.hpp): kernels of orders 17..24.
#include "mfma_prec.hpp"

namespace sfmm {
namespace mprec {

template <int... Ps>
constexpr auto make_fns_hi(std::integer_sequence<int, Ps...>)
{
    struct T {
        fn_t f[sizeof...(Ps)];
    } t{{&k_m2l_mfma_f32<Ps + 17>...}};
    return t;
}

fn_t kernel_hi(uint32_t P)
{
    static const auto fns = make_fns_hi(std::make_integer_sequence<int, kMax - 16>{});
    return (P >= 17 && P <= uint32_t(kMax)) ? fns.f[P - 17] : nullptr;
}

}  // namespace mprec
}  // namespace sfmm
