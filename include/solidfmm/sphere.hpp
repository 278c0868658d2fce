// Drop-in for solidfmm/sphere.hpp of vorticle/solidfmm 1.5 (sphere.hpp:28-42):
// a ball (centre and radius) and the smallest enclosing ball of a set of
// points or of a set of balls.  An empty set throws std::logic_error, as in
// the reference (sphere.cpp:551-553).
#ifndef SOLIDFMM_SPHERE_HPP
#define SOLIDFMM_SPHERE_HPP

#include <cstddef>

namespace solidfmm
{

template <typename real>
struct sphere
{
    real  x { 0 }, y { 0 }, z { 0 }; // centre
    real  r { 0 };                   // radius
};

extern template struct sphere<float >;
extern template struct sphere<double>;

sphere<float > bounding_sphere( size_t n, const float  *x, const float  *y, const float  *z );
sphere<double> bounding_sphere( size_t n, const double *x, const double *y, const double *z );

sphere<float > bounding_sphere( const sphere<float > *begin, const sphere<float > *end );
sphere<double> bounding_sphere( const sphere<double> *begin, const sphere<double> *end );

}

#endif
