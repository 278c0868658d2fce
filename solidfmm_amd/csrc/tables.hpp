// Compile-time operator tables for the translation kernels.
//
// All tables are evaluated by the compiler in IEEE double arithmetic, with
// the operation order of the reference's own table builders, so each entry
// is the same double the reference stores:
//   swap matrices B_n^{m,l}   swap_matrix.cpp:148-171 (Dehnen's recurrence)
//   folded real/imag blocks   operator_data.cpp:47-147
//   factorials                operator_data.cpp:195-196
//   inverse factorials        operator_data.cpp:203-214
// Kernels read them with compile-time indices, so every entry becomes an
// immediate operand (or is skipped when it is zero).
#pragma once

#include <type_traits>

namespace sfmm {

// Highest order (degrees 0..kMaxOrder-1) the kernels are instantiated for.
inline constexpr int kMaxOrder = 32;

template <int N>
struct SwapArr {
    double v[(2 * N + 1) * (2 * N + 1)];
};

template <int N>
constexpr SwapArr<N> make_swap();

template <int N>
inline constexpr SwapArr<N> kSwap = make_swap<N>();

// B_N^{m,l}; zero outside |m|,|l| <= N.
template <int N>
constexpr double swapB(int m, int l)
{
    if (m < -N || m > N || l < -N || l > N) return 0.0;
    return kSwap<N>.v[(m + N) + (l + N) * (2 * N + 1)];
}

template <int N>
constexpr SwapArr<N> make_swap()
{
    SwapArr<N> a{};
    if constexpr (N == 0) {
        a.v[0] = 1.0;
    } else {
        constexpr int D = 2 * N + 1;
        auto at = [](int m, int l) { return (m + N) + (l + N) * D; };
        for (int l = -N; l <= N; ++l) a.v[at(0, l)] = 0.5 * (swapB<N - 1>(0, l - 1) - swapB<N - 1>(0, l + 1));
        for (int m = 0; m < N; ++m)
            for (int l = -N; l <= N; ++l) {
                a.v[at(m + 1, l)] = 0.5 * (swapB<N - 1>(m, l - 1) + swapB<N - 1>(m, l + 1)) + swapB<N - 1>(m, l);
                a.v[at(-m - 1, l)] =
                    0.5 * (swapB<N - 1>(-m, l - 1) + swapB<N - 1>(-m, l + 1)) - swapB<N - 1>(-m, l);
            }
    }
    return a;
}

// Folded (N+1)x(N+1) blocks acting on the stored m >= 0 coefficients.
// Kind: 0 real, 1 imaginary (M-type pipeline); 2, 3 the transposed pair
// (L-type pipeline).  Row index is the output order m, column the input
// order l.  Real blocks vanish unless (l+m+N) is even, imaginary blocks
// unless it is odd.
template <int N, int KIND>
constexpr double folded(int m, int l)
{
    constexpr bool real_part = (KIND == 0 || KIND == 2);
    constexpr bool tr = (KIND >= 2);
    const bool odd = ((l + m + N) & 1) != 0;
    if (real_part ? odd : !odd) return 0.0;
    const double bp = tr ? swapB<N>(m, l) : swapB<N>(l, m);
    const double bm = tr ? swapB<N>(m, -l) : swapB<N>(-l, m);
    if (real_part) {
        if (l == 0) return tr ? swapB<N>(m, 0) : swapB<N>(0, m);
        return (l & 1) ? bp - bm : bp + bm;
    }
    return (l & 1) ? bp + bm : bp - bm;
}

// First column l whose entry of row m is non-zero, or -1.
template <int N, int KIND>
constexpr int first_nonzero(int m)
{
    for (int l = 0; l <= N; ++l)
        if (folded<N, KIND>(m, l) != 0.0) return l;
    return -1;
}

struct FacTable {
    double fac[2 * kMaxOrder];
    double inv[2 * kMaxOrder];
};

constexpr FacTable make_fac()
{
    FacTable t{};
    t.fac[0] = 1.0;
    for (int i = 1; i < 2 * kMaxOrder; ++i) t.fac[i] = static_cast<double>(i) * t.fac[i - 1];
    for (int i = 0; i < 2 * kMaxOrder; ++i) t.inv[i] = 1.0 / t.fac[i];
    return t;
}

inline constexpr FacTable kFac = make_fac();

// 2^k, exact.
constexpr double pow2(int k)
{
    double v = 1.0;
    for (; k > 0; --k) v *= 2.0;
    for (; k < 0; ++k) v *= 0.5;
    return v;
}

// floor(log2((n-m)! (n+m)!) / 2): 2^enorm(n, m) approximates the norm
// sqrt((n-m)! (n+m)!) that makes the swap blocks of degree n near-orthogonal
// (the reduced-precision kernels work in that scaled basis).
constexpr int enorm(int n, int m)
{
    double p = kFac.fac[n - m] * kFac.fac[n + m];
    int e = 0;
    while (p >= 2.0) {
        p *= 0.5;
        ++e;
    }
    return e >> 1;
}

template <int I>
using ic = std::integral_constant<int, I>;

// Compile-time loops: f(ic<B>), f(ic<B+1>), ..., f(ic<E-1>).
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

// Descending: f(ic<E-1>), ..., f(ic<B>).
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for_down(F &&f)
{
    if constexpr (B < E) {
        f(ic<E - 1>{});
        static_for_down<B, E - 1>(f);
    }
}

}  // namespace sfmm
