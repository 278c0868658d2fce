// Drop-in for solidfmm/threadlocal_buffer.hpp of vorticle/solidfmm 1.5
// (threadlocal_buffer.hpp:30-70): the per-thread workspace of the
// translation calls.  Here it holds the host staging the calls reuse (packed
// pools and item lists); one buffer per calling thread, as in the
// reference (doc/solidfmm.tex:1776-1824).
#ifndef SOLIDFMM_TRHEADLOCAL_BUFFER_HPP
#define SOLIDFMM_TRHEADLOCAL_BUFFER_HPP

#include <cstddef>

namespace solidfmm
{

template <typename real> class operator_data;

template <typename real>
class threadlocal_buffer
{
public:
    threadlocal_buffer() noexcept;
    threadlocal_buffer( const operator_data<real> &op );
    threadlocal_buffer( const threadlocal_buffer  &rhs );
    threadlocal_buffer(       threadlocal_buffer &&rhs ) noexcept;
    threadlocal_buffer& operator=( const threadlocal_buffer  &rhs );
    threadlocal_buffer& operator=(       threadlocal_buffer &&rhs ) noexcept;
   ~threadlocal_buffer();

    void swap( threadlocal_buffer &rhs ) noexcept;

    void initialise( const operator_data<real> &op );
    bool compatible( const operator_data<real> &op ) const noexcept;

    struct staging;                        // defined by the library
    staging* stage() noexcept { return s; }

private:
    size_t P { 0 };
    staging *s { nullptr };
};

extern template class threadlocal_buffer<float>;
extern template class threadlocal_buffer<double>;

}

#endif
