// Drop-in for solidfmm/operator_data.hpp of vorticle/solidfmm 1.5
// (operator_data.hpp:30-67).  The reference keeps its translation tables
// here; this class owns the device context (include/solidfmm_gpu.h) whose
// tables live on the GPU.  Orders: 1..86 (double), 1..18 (float), as in
// the reference (operator_data.cpp:35-45, 151-164).
#ifndef SOLIDFMM_OPERATOR_DATA_HPP
#define SOLIDFMM_OPERATOR_DATA_HPP

#include <cstddef>

extern "C" { struct sfmm_context; }

namespace solidfmm
{

template <typename real>
class operator_data
{
public:
    operator_data() = delete;
    operator_data( size_t P );   // logic_error if P == 0, overflow_error above the limit
    operator_data( const operator_data&  rhs ) = delete;
    operator_data(       operator_data&& rhs ) = delete;
    operator_data& operator=( const operator_data  &rhs ) = delete;
    operator_data& operator=(       operator_data &&rhs ) = delete;
    ~operator_data();

    size_t order() const noexcept { return P; }

    // The device context the translations run on (no counterpart in the
    // reference).
    sfmm_context* context() const noexcept { return ctx; }

private:
    size_t P;
    sfmm_context *ctx;
};

extern template class operator_data<float>;
extern template class operator_data<double>;

}

#endif
