// Instantiations of the one-lane-per-translation kernels for lists whose
// items all have the launch's orders (square orders 1..12).
#include <utility>

#include "lane_kernel.hpp"

namespace sfmm {
namespace lane {

const Table &fixed_table()
{
    static const Table t = make_table<false>();
    return t;
}

}  // namespace lane
}  // namespace sfmm
