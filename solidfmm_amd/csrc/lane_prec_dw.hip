// Double-word FP32 instantiations of the reduced-precision lane M2L kernel
// (orders 1..kPrecLaneMaxDW; higher orders use the generic precision
// kernel).
#include "lane_prec.hpp"

namespace sfmm {
namespace lprec {

kernel_fn kernel_dw(uint32_t P)
{
    static const auto t = make_table<prec::DW32>(std::make_integer_sequence<int, int(kPrecLaneMaxDW)>{});
    return (P >= 1 && P <= kPrecLaneMaxDW) ? t.f[P - 1] : nullptr;
}

}  // namespace lprec
}  // namespace sfmm
