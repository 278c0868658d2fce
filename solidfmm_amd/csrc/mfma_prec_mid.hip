// FP32 M2L on the matrix cores (mfma_prec.hpp): kernels of orders 17..20.
#include "mfma_prec.hpp"

namespace sfmm {
namespace mprec {

fn_t kernel_mid(uint32_t P)
{
    static const auto fns = make_fns<17>(std::make_integer_sequence<int, 4>{});
    return (P >= 17 && P <= 20) ? fns.f[P - 17] : nullptr;
}

}  // namespace mprec
}  // namespace sfmm
