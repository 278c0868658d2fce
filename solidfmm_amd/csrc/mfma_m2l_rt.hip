// Instantiations of the MFMA translation kernels for M2L (square orders
// kMfmaMin..kMfmaMax, runtime per-item orders).
#include "translate_mfma.hpp"

namespace sfmm {
namespace mfma {

template hipError_t launch_op<SFMM_M2L, true>(uint32_t, const sfmm_item *, size_t, const double *, double *,
                                                   const double *, hipStream_t, int);

}  // namespace mfma
}  // namespace sfmm
