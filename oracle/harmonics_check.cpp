// Test program for the public harmonics API (harmonics.hpp: R, S, dR, dS),
// compiled twice by the Makefile's dropin target: against the reference
// library and against the drop-in headers with libsolidfmm_gpu.so.  It
// prints every coefficient of R, S, dR and dS at fixed points and orders,
// in float and double; tests/test_harmonics_dropin.py compares the outputs.
#include <solidfmm/harmonics.hpp>
#include <solidfmm/solid.hpp>

#include <cstdio>
#include <initializer_list>

namespace {

template <typename real>
void dump(const char *what, const solidfmm::solid<real> &s)
{
    const size_t P = s.order(), dim = s.dimension();
    std::printf("%s %zu %zu %d", what, P, dim, int(sizeof(real)));
    for (size_t d = 0; d < dim; ++d)
        for (size_t n = 0; n < P; ++n)
            for (size_t m = 0; m <= n; ++m) std::printf(" %.17g %.17g", double(s.re(n, m, d)), double(s.im(n, m, d)));
    std::printf("\n");
}

template <typename real>
void run(size_t P, real x, real y, real z)
{
    dump("R", solidfmm::harmonics::R<real>(P, x, y, z));
    dump("S", solidfmm::harmonics::S<real>(P, x, y, z));
    dump("dR", solidfmm::harmonics::dR<real>(P, x, y, z));
    dump("dS", solidfmm::harmonics::dS<real>(P, x, y, z));
}

}  // namespace

int main()
{
    const double pts[][3] = {{0.3, -1.1, 0.8}, {1.0, 2.0, -0.5}, {0.0, 0.0, 1.7}, {-0.2, 0.05, 0.0}};
    for (size_t P : {1, 2, 5, 12}) {
        for (const auto &p : pts) {
            run<double>(P, p[0], p[1], p[2]);
            run<float>(P, float(p[0]), float(p[1]), float(p[2]));
        }
    }
    return 0;
}
