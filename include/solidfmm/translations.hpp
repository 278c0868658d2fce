// Drop-in for solidfmm/translations.hpp of vorticle/solidfmm 1.5
// (translations.hpp:28-175): the batch translator and the single-thread
// M2M / M2L / L2L entry points, with the reference's signatures, argument
// meaning and exceptions.  Every call runs on the GPU through the C ABI of
// solidfmm_gpu.h: the solids are staged into device pools, translated by the
// HIP kernels and accumulated back (target += T(source), items applied in
// list order per target).
//
// Exceptions, as in the reference (translations.cpp:634-1001): logic_error
// for an incompatible buffer or a dimension mismatch, out_of_range for an
// order above the operator data's order; everything is checked before any
// target is written.  The *_unchecked variants skip the checks.
#ifndef SOLIDFMM_TRANSLATIONS_HPP
#define SOLIDFMM_TRANSLATIONS_HPP

#include <cstddef>

extern "C" { struct hwloc_topology; struct hwloc_bitmap_s; }

namespace solidfmm
{

template <typename real> class solid;

//////////////////////////////
// Multithreaded interface. //
//////////////////////////////

template <typename real>
struct translation_info
{
    const solid<real> *source;
          solid<real> *target;
    real x, y, z;
};

template <typename real>
class multithreaded_translator
{
public:
    using item = translation_info<real>;
    using cpuset_t = hwloc_bitmap_s*;

    multithreaded_translator() = delete;
    multithreaded_translator( const multithreaded_translator  &rhs ) = delete;
    multithreaded_translator(       multithreaded_translator &&rhs ) noexcept;
    ~multithreaded_translator();
    multithreaded_translator& operator=( const multithreaded_translator  &rhs ) = delete;
    multithreaded_translator& operator=(       multithreaded_translator &&rhs ) noexcept;

    // Runs on the first visible GPU.
    multithreaded_translator( size_t max_expansion_order );

    // Accepted for source compatibility; the work runs on the GPU, so the
    // topology and cpusets are not used.
    multithreaded_translator( size_t max_expansion_order,
                              hwloc_topology *topology,
                              cpuset_t *begin, cpuset_t *end );

    // Scheduling policy of the CPU thread pool in the reference; no effect
    // here (one device queue).
    void schedule_equally_among_threads() noexcept;
    void schedule_by_measuring_throughput();

    void m2m( const item *begin, const item *end ) const;
    void m2l( const item *begin, const item *end ) const;
    void l2l( const item *begin, const item *end ) const;

    void m2m_unchecked( const item *begin, const item *end ) const noexcept;
    void m2l_unchecked( const item *begin, const item *end ) const noexcept;
    void l2l_unchecked( const item *begin, const item *end ) const noexcept;

private:
    class impl;
    impl* p;
};

extern template class multithreaded_translator<float>;
extern template class multithreaded_translator<double>;

////////////////////////////////
// Single threaded functions. //
////////////////////////////////

template <typename real> class operator_data;
template <typename real> class threadlocal_buffer;

void m2l( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          size_t howmany, const solid<float> *const *const M, solid<float> *const *const L,
          const float *x, const float *y, const float *z );
void m2l( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          size_t howmany, const solid<double> *const *const M, solid<double> *const *const L,
          const double *x, const double *y, const double *z );
void m2l( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          const translation_info<float> *begin, const translation_info<float> *end );
void m2l( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          const translation_info<double> *begin, const translation_info<double> *end );

void m2m( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          size_t howmany, const solid<float> *const *const Min, solid<float> *const *const Mout,
          const float *x, const float *y, const float *z );
void m2m( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          size_t howmany, const solid<double> *const *const Min, solid<double> *const *const Mout,
          const double *x, const double *y, const double *z );
void m2m( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          const translation_info<float> *begin, const translation_info<float> *end );
void m2m( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          const translation_info<double> *begin, const translation_info<double> *end );

void l2l( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          size_t howmany, const solid<float> *const *const Min, solid<float> *const *const Mout,
          const float *x, const float *y, const float *z );
void l2l( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          size_t howmany, const solid<double> *const *const Lin, solid<double> *const *const Lout,
          const double *x, const double *y, const double *z );
void l2l( const operator_data<float> &op, threadlocal_buffer<float> &buf,
          const translation_info<float> *begin, const translation_info<float> *end );
void l2l( const operator_data<double> &op, threadlocal_buffer<double> &buf,
          const translation_info<double> *begin, const translation_info<double> *end );

void m2l_unchecked( const operator_data<float> &op, threadlocal_buffer<float> &buf,
                    const translation_info<float> *begin, const translation_info<float> *end ) noexcept;
void m2l_unchecked( const operator_data<double> &op, threadlocal_buffer<double> &buf,
                    const translation_info<double> *begin, const translation_info<double> *end ) noexcept;
void m2m_unchecked( const operator_data<float> &op, threadlocal_buffer<float> &buf,
                    const translation_info<float> *begin, const translation_info<float> *end ) noexcept;
void m2m_unchecked( const operator_data<double> &op, threadlocal_buffer<double> &buf,
                    const translation_info<double> *begin, const translation_info<double> *end ) noexcept;
void l2l_unchecked( const operator_data<float> &op, threadlocal_buffer<float> &buf,
                    const translation_info<float> *begin, const translation_info<float> *end ) noexcept;
void l2l_unchecked( const operator_data<double> &op, threadlocal_buffer<double> &buf,
                    const translation_info<double> *begin, const translation_info<double> *end ) noexcept;

}

#endif
