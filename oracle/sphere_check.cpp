// Test program for the public sphere API (sphere.hpp:38-42), compiled twice
// by the Makefile's dropin target: against the reference library and
// against the drop-in headers with libsolidfmm_gpu.so.  It prints the
// smallest balls of fixed pseudo-random point and ball sets, in float and
// double, one line per set; tests/test_sphere.py compares the two outputs.
#include <solidfmm/sphere.hpp>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace {

uint64_t state = 0x5011DF0Bu;
double next01()
{
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return double(state >> 11) * (1.0 / 9007199254740992.0);
}

template <typename real>
void points(size_t n)
{
    std::vector<real> x(n), y(n), z(n);
    for (size_t i = 0; i < n; ++i) {
        x[i] = real(2 * next01() - 1);
        y[i] = real(2 * next01() - 1);
        z[i] = real(next01());
    }
    const solidfmm::sphere<real> s = solidfmm::bounding_sphere(n, x.data(), y.data(), z.data());
    std::printf("points %zu %d %.17g %.17g %.17g %.17g\n", n, int(sizeof(real)), double(s.x), double(s.y),
                double(s.z), double(s.r));
}

template <typename real>
void balls(size_t n)
{
    std::vector<solidfmm::sphere<real>> b(n);
    for (auto &s : b) {
        s.x = real(2 * next01() - 1);
        s.y = real(2 * next01() - 1);
        s.z = real(2 * next01() - 1);
        s.r = real(0.5 * next01());
    }
    const solidfmm::sphere<real> s = solidfmm::bounding_sphere(b.data(), b.data() + n);
    std::printf("balls %zu %d %.17g %.17g %.17g %.17g\n", n, int(sizeof(real)), double(s.x), double(s.y),
                double(s.z), double(s.r));
}

}  // namespace

int main()
{
    for (size_t n : {1, 2, 3, 4, 5, 10, 100, 1000}) {
        points<double>(n);
        points<float>(n);
        balls<double>(n);
        balls<float>(n);
    }
    try {
        solidfmm::bounding_sphere(0, static_cast<const double *>(nullptr), nullptr, nullptr);
        std::printf("empty: no exception\n");
    } catch (const std::logic_error &) {
        std::printf("empty: logic_error\n");
    }
    return 0;
}
