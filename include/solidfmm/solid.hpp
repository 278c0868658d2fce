// Drop-in for solidfmm/solid.hpp of vorticle/solidfmm 1.5 (solid.hpp:28-79):
// the coefficient container of an order-P solid-harmonic expansion and the
// helpers dot() and fmadd() (solid.cpp:40-105).  Same class layout rules:
// Re C_n^m at n(n+1)+m, Im C_n^m at (n+1)^2+m, P(P+1) reals per dimension.
#ifndef SOLIDFMM_SOLID_HPP
#define SOLIDFMM_SOLID_HPP

#include <cstddef>

namespace solidfmm
{

template <typename real>
class solid
{
public:
    solid() = default;
    solid( size_t P, size_t dim = 1 );
    solid( const solid  &rhs );
    solid(       solid &&rhs ) noexcept;
    solid& operator=( const solid  &rhs );
    solid& operator=(       solid &&rhs ) noexcept;
   ~solid();

    void resize( size_t P, size_t dim = 1 );   // storage for (P, dim); contents not remapped
    void reinit( size_t P, size_t dim = 1 );   // resize, then zeros()
    void zeros() noexcept;

    const real& re( size_t n, size_t m, size_t d = 0 ) const noexcept { return data[ ridx(n,m,d) ]; }
          real& re( size_t n, size_t m, size_t d = 0 )       noexcept { return data[ ridx(n,m,d) ]; }

    const real& im( size_t n, size_t m, size_t d = 0 ) const noexcept { return data[ iidx(n,m,d) ]; }
          real& im( size_t n, size_t m, size_t d = 0 )       noexcept { return data[ iidx(n,m,d) ]; }

    size_t       order    () const noexcept { return P;   }
    size_t       dimension() const noexcept { return dim; }

    const real*  memptr() const noexcept { return data; }
          real*  memptr()       noexcept { return data; }

private:
    size_t ridx( size_t n, size_t m, size_t d ) const noexcept { return d*P*(P+1) + n*(n+1) + m; }
    size_t iidx( size_t n, size_t m, size_t d ) const noexcept { return d*P*(P+1) + (n+1)*(n+1) + m; }

    size_t  P     { 0 };
    size_t  dim   { 1 };
    real   *data  { nullptr };
};

extern template class solid<float>;
extern template class solid<double>;

void dot( const solid<float > &A, const solid<float > &B, float  *result ) noexcept;
void dot( const solid<double> &A, const solid<double> &B, double *result ) noexcept;

void fmadd( const float   fac, const solid<float > &A, solid<float > &B );
void fmadd( const float  *fac, const solid<float > &A, solid<float > &B );
void fmadd( const double  fac, const solid<double> &A, solid<double> &B );
void fmadd( const double *fac, const solid<double> &A, solid<double> &B );

}

#endif
