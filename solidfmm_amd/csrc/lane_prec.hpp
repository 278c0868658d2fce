// Reduced-precision M2L, one lane per translation, square orders 1..16
// (FP32; double-word FP32 up to 10):
// the register-resident mapping of lane_kernel.hpp with the bulk arithmetic
// (rotations, swaps, z-translation) in an arithmetic policy of prec.hpp.
//
// Scaled basis.  The folded swap blocks of solidfmm's normalisation have
// entries up to ~360 at degree 9 and ~2e6 at degree 23
// (operator_data.cpp:47-147), which costs FP32 digits to cancellation and,
// at high orders, its exponent range.  Order m of degree n is therefore
// carried scaled by 2^(e(n,m) - c_n), with 2^e(n,m) ~ sqrt((n-m)!(n+m)!)
// (tables.hpp enorm): the M-type blocks become 2^e(n,m) S 2^-e(n,l) and the
// L-type (transposed) blocks 2^-e(n,m) S 2^e(n,l), both near-orthogonal.  All
// these factors are powers of two, so the change of basis is exact.
//
// FP64 is kept for the geometry and the scaling: Euler powers are computed
// in FP64 and converted once; each source degree K is scaled by
// K! r^-(K+1) 2^K in FP64 and converted; the z-translation weight
// (N+K)! (translations.cpp:283-296) becomes C(N+K, K) 2^-(N+K) times the
// basis factors, and the backward scale N! r^-N 2^N is applied in FP64 to
// the converted result just before the FP64 accumulation into the target.
// Source and target degrees are fetched one degree ahead of their use.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "../../include/solidfmm_gpu.h"
#include "prec.hpp"
#include "tables.hpp"
#include "translate.hpp"

namespace sfmm {
namespace lprec {

constexpr int tri(int m, int n) { return n * (n + 1) / 2 + m; }

constexpr double binom(int n, int k)
{
    double c = 1.0;
    for (int j = 1; j <= k; ++j) c = c * double(n - k + j) / double(j);
    return c;
}

__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// Basis exponents of order m, degree n: forward M-type (source side) and
// backward L-type (target side) scalings.
constexpr int qf(int n, int m) { return enorm(n, m) - enorm(n, n); }
constexpr int qb(int n, int m) { return enorm(n, 0) - enorm(n, m); }

// Swap of degree N in the scaled basis: blocks KR / KI (0, 1: M-type with
// exponents qf; 2, 3: L-type with exponents qb).
template <class Pol, int N, int KR, int KI>
__device__ __forceinline__ void swap_stage(typename Pol::T (&re)[N + 1], typename Pol::T (&im)[N + 1])
{
    using T = typename Pol::T;
    constexpr bool ltype = KR >= 2;
    T ore[N + 1], oim[N + 1];
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int fr = first_nonzero<N, KR>(m);
        constexpr int fi = first_nonzero<N, KI>(m);
        T ar = Pol::zero(), ai = Pol::zero();
        static_for<0, N + 1>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            constexpr double sc = ltype ? pow2(qb(N, m) - qb(N, l)) : pow2(qf(N, m) - qf(N, l));
            constexpr double cr = folded<N, KR>(m, l) * sc;
            constexpr double ci = folded<N, KI>(m, l) * sc;
            if constexpr (cr != 0.0) {
                if constexpr (l == fr) ar = Pol::mul(Pol::from(cr), re[l]);
                else ar = Pol::fma(Pol::from(cr), re[l], ar);
            }
            if constexpr (ci != 0.0) {
                if constexpr (l == fi) ai = Pol::mul(Pol::from(ci), im[l]);
                else ai = Pol::fma(Pol::from(ci), im[l], ai);
            }
        });
        ore[m] = ar;
        oim[m] = ai;
    });
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        re[m] = ore[m];
        im[m] = oim[m];
    });
}

// (re,im) <- e^{+i m theta} (re,im), or e^{-i m theta} when !FWD.
template <class Pol, int N, bool FWD>
__device__ __forceinline__ void rot(const typename Pol::T *c, const typename Pol::T *s, typename Pol::T (&re)[N + 1],
                                    typename Pol::T (&im)[N + 1])
{
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const auto a = re[m], b = im[m];
        if constexpr (FWD) {
            re[m] = Pol::sub(Pol::mul(c[m], a), Pol::mul(s[m], b));
            im[m] = Pol::add(Pol::mul(s[m], a), Pol::mul(c[m], b));
        } else {
            re[m] = Pol::add(Pol::mul(c[m], a), Pol::mul(s[m], b));
            im[m] = Pol::sub(Pol::mul(c[m], b), Pol::mul(s[m], a));
        }
    });
}

template <int N, int CAP>
__device__ __forceinline__ void fetch_degree(const double *__restrict__ blk, double2 (&v)[CAP])
{
    const double2 *q = reinterpret_cast<const double2 *>(blk + N * (N + 1));
    static_for<0, N + 1>([&](auto jc) { v[decltype(jc)::value] = q[decltype(jc)::value]; });
}

// Zeros read in place of a target that is not accumulated (acc == 0): the
// loads are never skipped, so no branch splits the load stream.
inline __device__ double kZeroTarget[kMaxOrder * (kMaxOrder + 1)];

template <int N, int CAP>
__device__ __forceinline__ double piece(const double2 (&v)[CAP], int k)
{
    return (k & 1) ? v[k / 2].y : v[k / 2].x;
}

template <class Pol, int P>
__global__ void __launch_bounds__(64)
    k_m2l_prec(const double *__restrict__ src_pool, double *__restrict__ dst_pool, const sfmm_item *__restrict__ items,
               uint32_t n_items, int acc)
{
    using T = typename Pol::T;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_items) return;
    const sfmm_item it = items[t];
    const double *__restrict__ S = static_cast<const double *>(__builtin_assume_aligned(src_pool + it.src, 16));
    double *__restrict__ D = static_cast<double *>(__builtin_assume_aligned(dst_pool + it.dst, 16));
    const double *__restrict__ Dr = acc ? D : kZeroTarget;  // old target values
    const double x = it.x, y = it.y, z = it.z;

    // Euler data (microkernel_generic.cpp:61-96) in FP64, powers converted.
    const double rho = sqrt(x * x + y * y);
    const double r = sqrt(x * x + y * y + z * z);
    const double rinv = 1.0 / r;
    double c1 = 1.0, s1 = 0.0;
    if (rho != 0.0) {
        const double q = 1.0 / rho;
        c1 = y * q;
        s1 = x * q;
    }
    const double cb1 = z * rinv, sb1 = -rho * rinv;
    T ca[P], sa[P], cb[P], sb[P];
    {
        double pc = 1.0, ps = 0.0, qc = 1.0, qs = 0.0;
        static_for<0, P>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            ca[k] = Pol::from(pc);
            sa[k] = Pol::from(ps);
            cb[k] = Pol::from(qc);
            sb[k] = Pol::from(qs);
            const double a = pc * c1 - ps * s1;
            ps = ps * c1 + pc * s1;
            pc = a;
            const double b = qc * cb1 - qs * sb1;
            qs = qs * cb1 + qc * sb1;
            qc = b;
        });
    }

    T tre[tri(0, P)], tim[tri(0, P)];
    double2 sv[P], tg[P];
    fetch_degree<0>(S, sv);
    sched_fence();
    double fs = rinv;  // n! r^-(n+1) 2^n
    static_for<0, P>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        double2 cur[P];
        static_for<0, n + 1>([&](auto jc) { cur[decltype(jc)::value] = sv[decltype(jc)::value]; });
        if constexpr (n + 1 < P) fetch_degree<n + 1>(S, sv);
        else fetch_degree<P - 1>(Dr, tg);
        sched_fence();
        T re[n + 1], im[n + 1];
        static_for<0, n + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr double q = pow2(qf(n, m));
            re[m] = Pol::from(fs * piece<n>(cur, m) * q);
            im[m] = Pol::from(fs * piece<n>(cur, n + 1 + m) * q);
        });
        rot<Pol, n, true>(ca, sa, re, im);
        swap_stage<Pol, n, 0, 1>(re, im);
        rot<Pol, n, true>(cb, sb, re, im);
        swap_stage<Pol, n, 0, 1>(re, im);
        static_for<0, n + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            tre[tri(m, n)] = re[m];
            tim[tri(m, n)] = im[m];
        });
        fs *= rinv * double(2 * (n + 1));
    });

    // backward scales N! r^-N 2^N
    double bs[P];
    bs[0] = 1.0;
    static_for<1, P>([&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        bs[N] = bs[N - 1] * rinv * double(2 * N);
    });
    static_for_down<0, P>([&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        double2 old[P];
        static_for<0, N + 1>([&](auto jc) { old[decltype(jc)::value] = tg[decltype(jc)::value]; });
        if constexpr (N > 0) fetch_degree<N - 1>(Dr, tg);
        sched_fence();
        T re[N + 1], im[N + 1];
        static_for<0, N + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            T ar = Pol::zero(), ai = Pol::zero();
            static_for<m, P>([&](auto Kc) {
                constexpr int K = decltype(Kc)::value;
                constexpr double w = binom(N + K, K) * pow2(qb(N, m) - qf(K, m) - N - K);
                ar = Pol::fma(Pol::from(w), tre[tri(m, K)], ar);
                ai = Pol::fma(Pol::from(w), tim[tri(m, K)], ai);
            });
            if constexpr ((N + m) & 1) {
                ar = Pol::neg(ar);
                ai = Pol::neg(ai);
            }
            re[m] = ar;
            im[m] = ai;
        });
        swap_stage<Pol, N, 2, 3>(re, im);
        rot<Pol, N, false>(cb, sb, re, im);
        swap_stage<Pol, N, 2, 3>(re, im);
        rot<Pol, N, false>(ca, sa, re, im);
        double2 *q = reinterpret_cast<double2 *>(D + N * (N + 1));
        double out[2 * (N + 1)];
        static_for<0, N + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            constexpr double q = pow2(-qb(N, m));
            out[m] = bs[N] * (Pol::to(re[m]) * q);
            out[N + 1 + m] = bs[N] * (Pol::to(im[m]) * q);
        });
        static_for<0, N + 1>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            double2 v = old[j];
            v.x += out[2 * j];
            v.y += out[2 * j + 1];
            q[j] = v;
        });
    });
}

using kernel_fn = void (*)(const double *, double *, const sfmm_item *, uint32_t, int);

template <class Pol, int... Ps>
constexpr auto make_table(std::integer_sequence<int, Ps...>)
{
    struct T {
        kernel_fn f[sizeof...(Ps)];
    } t{{&k_m2l_prec<Pol, Ps + 1>...}};
    return t;
}

// Kernel of order P (1 <= P <= the policy's lane limit) per policy, each
// instantiated in a translation unit of its own (lane_prec_f32.hip,
// lane_prec_dw.hip).
kernel_fn kernel_f32(uint32_t P);
kernel_fn kernel_dw(uint32_t P);

}  // namespace lprec
}  // namespace sfmm
