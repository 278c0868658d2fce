// Instantiations of the one-lane-per-translation kernels for mixed-order
// lists: each item brings its own orders, bounded by the kernel's order.
#include <utility>

#include "lane_kernel.hpp"

namespace sfmm {
namespace lane {

const Table &rt_table()
{
    static const Table t = make_table<true>();
    return t;
}

}  // namespace lane
}  // namespace sfmm
