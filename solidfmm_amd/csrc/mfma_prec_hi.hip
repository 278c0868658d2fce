// FP32 M2L on the matrix cores (mfma_prec.hpp): kernels of orders 21..24.
#include "mfma_prec.hpp"

namespace sfmm {
namespace mprec {

fn_t kernel_hi(uint32_t P)
{
    static const auto fns = make_fns<21>(std::make_integer_sequence<int, kMax - 20>{});
    return (P >= 21 && P <= uint32_t(kMax)) ? fns.f[P - 21] : nullptr;
}

}  // namespace mprec
}  // namespace sfmm
