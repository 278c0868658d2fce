// FP32 M2L on the matrix cores: square orders 1..24, swaps on
// v_mfma_f32_16x16x4f32 with FP32 accumulation, the rotations and the
// z-translation on the FP32 VALU, geometry and the accumulation into the
// FP64 target in FP64.  Algorithm: Dehnen's rotation scheme as the
// reference implements it (translations.cpp:36-85 forward, :138-187
// backward, :283-296 z-translation), in the scaled basis of lane_prec.hpp.
//
// Mapping.  A wave owns 16 translations; lane l = t + 16 q handles
// translation t and the orders m = 16 b + 4 q + i of every degree ("slot"
// s = 4 b + i, block b = 0, 1).  That is how the FP32 MFMA lays out its
// 16x16 result (element i of lane (q, t) is row 4 q + i, column t); the
// K-chunk c of input block b' takes, from quarter-wave k, the slot 4 b' + c,
// i.e. input order 16 b' + 4 k + c, so the result registers of one swap are
// the B operands of the next with no shuffles, and the A operand of chunk
// c is the column set {16 b' + 4 k + c} of the folded block.
//
// Scaled basis (lane_prec.hpp): order m of degree n is carried scaled by
// 2^qf(n, m) through the forward swaps (M-type blocks 2^qf S 2^-qf) and by
// 2^qb(n, m) through the backward ones (L-type blocks 2^qb S^T' 2^-qb); all
// factors are powers of two, so the change of basis is exact.  The rotated
// triangle is kept unscaled, so the z-translation weight of (N, K) is
// C(N+K, K) 2^-(N+K) for every order, an immediate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "../../include/solidfmm_gpu.h"
#include "lane_prec.hpp"
#include "tables.hpp"
#include "translate.hpp"

namespace sfmm {
namespace mprec {

typedef float flt4 __attribute__((ext_vector_type(4)));

inline constexpr int kWaves = 4;          // waves per workgroup
inline constexpr int kMax = kPrecMaxOrder;  // orders 1..24
static_assert(kMax <= 32, "two 16-order blocks per lane");

// Dense FP32 blocks per degree n: M-type real, M-type imaginary, L-type
// real, L-type imaginary, each (n+1) rows (output order m) at an odd row
// stride (input order l).
constexpr int st32(int n) { return (n + 1) | 1; }
constexpr int part32(int n) { return (n + 1) * st32(n); }
constexpr int base32(int n)
{
    int s = 0;
    for (int j = 0; j < n; ++j) s += 4 * part32(j);
    return s;
}
// Powers of two of the basis, [kind][n][m] at kind * kMax^2 + n * kMax + m:
// kind 0 2^qf, 1 2^-qf, 2 2^qb, 3 2^-qb.
constexpr int kPow = 4 * kMax * kMax;

struct Tab32 {
    float swap[base32(kMax)];
    float pw[kPow];
};

template <int N>
constexpr void fill_degree(Tab32 &t)
{
    if constexpr (N < kMax) {
        constexpr int st = st32(N), pt = part32(N), b0 = base32(N);
        for (int m = 0; m <= N; ++m) {
            for (int l = 0; l <= N; ++l) {
                const double sm = pow2(lprec::qf(N, m) - lprec::qf(N, l));
                const double sl = pow2(lprec::qb(N, m) - lprec::qb(N, l));
                t.swap[b0 + m * st + l] = static_cast<float>(folded<N, 0>(m, l) * sm);
                t.swap[b0 + pt + m * st + l] = static_cast<float>(folded<N, 1>(m, l) * sm);
                t.swap[b0 + 2 * pt + m * st + l] = static_cast<float>(folded<N, 2>(m, l) * sl);
                t.swap[b0 + 3 * pt + m * st + l] = static_cast<float>(folded<N, 3>(m, l) * sl);
            }
            t.pw[0 * kMax * kMax + N * kMax + m] = static_cast<float>(pow2(lprec::qf(N, m)));
            t.pw[1 * kMax * kMax + N * kMax + m] = static_cast<float>(pow2(-lprec::qf(N, m)));
            t.pw[2 * kMax * kMax + N * kMax + m] = static_cast<float>(pow2(lprec::qb(N, m)));
            t.pw[3 * kMax * kMax + N * kMax + m] = static_cast<float>(pow2(-lprec::qb(N, m)));
        }
        fill_degree<N + 1>(t);
    }
}

constexpr Tab32 make_tab32()
{
    Tab32 t{};
    fill_degree<0>(t);
    return t;
}

// One copy per translation unit that instantiates kernels.
static __device__ constexpr Tab32 kTab32 = make_tab32();

// Read in place of a coefficient a lane does not load.
inline __device__ double kZeros[2];

__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

__device__ __forceinline__ flt4 mfma32(float a, float b, flt4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Slot s of a lane holds orders 16 (s / 4) + 4 q + (s % 4); kmin is the
// smallest of them (q = 0).
constexpr int kmin(int s) { return 16 * (s / 4) + (s % 4); }
template <int P>
constexpr int nslots()
{
    return 4 * ((P + 15) / 16);
}
// Rotated triangle: slot s keeps degrees kmin(s)..P-1.
template <int P>
constexpr int toff(int s)
{
    int o = 0;
    for (int j = 0; j < s; ++j) o += kmin(j) < P ? P - kmin(j) : 0;
    return o;
}
template <int P>
constexpr int tsize()
{
    return toff<P>(nslots<P>());
}

// One swap of degree N on the lane's slots: blocks PART (0: M-type pair,
// 2: L-type pair) of the LDS table.  Orders above N come out zero.
template <int N, int PART, int NSL>
__device__ __forceinline__ void swap32(const float *tab, int lane, float (&xr)[NSL], float (&xi)[NSL])
{
    constexpr int NB = N / 16 + 1;  // row blocks holding orders <= N
    constexpr int ST = st32(N), PT = part32(N);
    const float *sr = tab + base32(N) + PART * PT;
    const float *si = sr + PT;
    // recomputed per swap: indices kept from one degree to the next would
    // hold registers through the whole translation
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int r = ln & 15;  // A row within the block
    const int k = ln >> 4;  // A column within the chunk
    flt4 dr[NB], di[NB];
    static_for<0, NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        dr[b] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
        di[b] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
        const int m = 16 * b + r;
        static_for<0, NB>([&](auto pc) {
            constexpr int bp = decltype(pc)::value;
            static_for<0, 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if constexpr (16 * bp + c <= N) {
                    const int l = 16 * bp + 4 * k + c;
                    const bool in = m <= N && l <= N;
                    const int safe = in ? m * ST + l : 0;
                    // always read a valid entry, then mask by a factor (a
                    // select would become a branch around the LDS read)
                    const float fin = in ? 1.0f : 0.0f;
                    dr[b] = mfma32(sr[safe] * fin, xr[4 * bp + c], dr[b]);
                    di[b] = mfma32(si[safe] * fin, xi[4 * bp + c], di[b]);
                }
            });
        });
    });
    static_for<0, NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        static_for<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            xr[4 * b + i] = dr[b][i];
            xi[4 * b + i] = di[b][i];
        });
    });
}

// (re, im) <- e^{+i m theta} (re, im) (FWD) or e^{-i m theta} (re, im) on
// the slots that hold orders <= N.
template <int N, bool FWD, int NSL>
__device__ __forceinline__ void rot(const float (&c)[NSL], const float (&s)[NSL], float (&xr)[NSL], float (&xi)[NSL])
{
    static_for<0, NSL>([&](auto sc) {
        constexpr int j = decltype(sc)::value;
        if constexpr (kmin(j) <= N) {
            const float a = xr[j], b = xi[j];
            if constexpr (FWD) {
                xr[j] = c[j] * a - s[j] * b;
                xi[j] = s[j] * a + c[j] * b;
            } else {
                xr[j] = c[j] * a + s[j] * b;
                xi[j] = c[j] * b - s[j] * a;
            }
        }
    });
}

template <int P>
__device__ __forceinline__ void m2l_group(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                                          const sfmm_item *__restrict__ items, uint32_t n_items, uint32_t g,
                                          const float *tab, const float *pw, int lane, bool acc)
{
    constexpr int NSL = nslots<P>();
    const int q = lane >> 4;
    const int t = lane & 15;
    const uint32_t idx = g * 16u + static_cast<uint32_t>(t);
    const bool valid = idx < n_items;

    // Every lane stays active to the end: the MFMAs need the whole wave.
    sfmm_item it{};
    if (valid) it = items[idx];
    const double *__restrict__ S = src_pool + it.src;
    double *__restrict__ D = dst_pool + it.dst;
    double x = it.x, y = it.y, z = it.z;
    if (!valid) {  // harmless geometry for idle lanes
        x = 0.0;
        y = 0.0;
        z = 1.0;
    }

    // Euler data (microkernel_generic.cpp:61-96) in FP64; the powers
    // e^{i k theta}, k = 16 b + 4 q + i, are converted once.
    const double rho = sqrt(x * x + y * y);
    const double r = sqrt(x * x + y * y + z * z);
    const double rinv = 1.0 / r;
    double ca1 = 1.0, sa1 = 0.0;
    if (rho != 0.0) {
        const double qq = 1.0 / rho;
        ca1 = y * qq;
        sa1 = x * qq;
    }
    const double cb1 = z * rinv, sb1 = -rho * rinv;
    float ca[NSL], sa[NSL], cb[NSL], sb[NSL];
    {
        double pc = 1.0, ps = 0.0, qc = 1.0, qs = 0.0;
        static_for<0, 4 * NSL>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int slot = 4 * (k / 16) + (k % 4);
            const bool mine = ((k % 16) / 4) == q;  // selects, not branches
            ca[slot] = mine ? static_cast<float>(pc) : ca[slot];
            sa[slot] = mine ? static_cast<float>(ps) : sa[slot];
            cb[slot] = mine ? static_cast<float>(qc) : cb[slot];
            sb[slot] = mine ? static_cast<float>(qs) : sb[slot];
            if constexpr (k + 1 < 4 * NSL) {
                const double a = pc * ca1 - ps * sa1;
                ps = ps * ca1 + pc * sa1;
                pc = a;
                const double b = qc * cb1 - qs * sb1;
                qs = qs * cb1 + qc * sb1;
                qc = b;
            }
        });
    }
    auto order = [&](int s) { return 16 * (s / 4) + 4 * q + (s % 4); };

    // Source degree n of the lane's orders, FP64 (zeros above n).
    auto load_degree = [&](auto nc, double(&vr)[NSL], double(&vi)[NSL]) {
        constexpr int n = decltype(nc)::value;
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            vr[s] = 0.0;
            vi[s] = 0.0;
            if constexpr (kmin(s) <= n) {
                const int m = order(s);
                const bool ok = valid && m <= n;
                const double *p = ok ? S + (n * (n + 1) + m) : kZeros;
                vr[s] = p[0];
                vi[s] = p[ok ? n + 1 : 0];  // Im C_n^m follows Re C_n^m by n+1
            }
        });
    };
    // Target degree N (zeros when not accumulating).
    auto fetch_old = [&](auto Nc, double(&vr)[NSL], double(&vi)[NSL]) {
        constexpr int N = decltype(Nc)::value;
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            vr[s] = 0.0;
            vi[s] = 0.0;
            if constexpr (kmin(s) <= N) {
                const int m = order(s);
                const bool ok = valid && acc && m <= N;
                const double *p = ok ? D + (N * (N + 1) + m) : kZeros;
                vr[s] = p[0];
                vi[s] = p[ok ? N + 1 : 0];
            }
        });
    };

    float tr[tsize<P>()], ti[tsize<P>()];
    double pr[NSL], pim[NSL];  // prefetched source degree
    double orr[NSL], oi[NSL];  // prefetched target degree

    // Forward, all source degrees: scale n! r^-(n+1) 2^n 2^qf(n, m) in FP64,
    // rotate, swap, rotate, swap in FP32, keep unscaled (2^-qf).
    double fs = rinv;
    load_degree(ic<0>{}, pr, pim);
    static_for<0, P>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        double vr[NSL], vi[NSL];
        static_for<0, NSL>([&](auto sc) {
            vr[decltype(sc)::value] = pr[decltype(sc)::value];
            vi[decltype(sc)::value] = pim[decltype(sc)::value];
        });
        if constexpr (n + 1 < P) load_degree(ic<n + 1>{}, pr, pim);
        else fetch_old(ic<0>{}, orr, oi);
        sched_fence();
        float xr[NSL], xi[NSL];
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            xr[s] = 0.0f;
            xi[s] = 0.0f;
            if constexpr (kmin(s) <= n) {
                const int m = order(s);
                const int mc = m <= n ? m : 0;
                const double f = fs * static_cast<double>(pw[0 * kMax * kMax + n * kMax + mc]);
                xr[s] = static_cast<float>(vr[s] * f);
                xi[s] = static_cast<float>(vi[s] * f);
            }
        });
        rot<n, true>(ca, sa, xr, xi);
        swap32<n, 0>(tab, lane, xr, xi);
        rot<n, true>(cb, sb, xr, xi);
        swap32<n, 0>(tab, lane, xr, xi);
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (kmin(s) <= n) {
                const int m = order(s);
                const int mc = m <= n ? m : 0;
                const float u = pw[1 * kMax * kMax + n * kMax + mc];
                tr[toff<P>(s) + n - kmin(s)] = xr[s] * u;
                ti[toff<P>(s) + n - kmin(s)] = xi[s] * u;
            }
        });
        fs *= rinv * static_cast<double>(2 * (n + 1));
    });

    // Output degrees ascending: L_N^m = (-1)^(N+m) sum_K (N+K)! T_K^m,
    // here C(N+K, K) 2^-(N+K) on the unscaled triangle, then 2^qb(N, m) into
    // the backward basis; swap, rotate back, swap, rotate back; N! r^-N 2^N
    // 2^-qb(N, m) in FP64 and the FP64 accumulation into the target.
    double bs = 1.0;
    static_for<0, P>([&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        double cr[NSL], ci[NSL];
        static_for<0, NSL>([&](auto sc) {
            cr[decltype(sc)::value] = orr[decltype(sc)::value];
            ci[decltype(sc)::value] = oi[decltype(sc)::value];
        });
        if constexpr (N + 1 < P) fetch_old(ic<N + 1>{}, orr, oi);
        sched_fence();
        float xr[NSL], xi[NSL];
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            xr[s] = 0.0f;
            xi[s] = 0.0f;
            if constexpr (kmin(s) <= N) {
                float ar = 0.0f, ai = 0.0f;
                static_for<kmin(s), P>([&](auto Kc) {
                    constexpr int K = decltype(Kc)::value;
                    constexpr float w = static_cast<float>(lprec::binom(N + K, K) * pow2(-N - K));
                    ar = __builtin_fmaf(w, tr[toff<P>(s) + K - kmin(s)], ar);
                    ai = __builtin_fmaf(w, ti[toff<P>(s) + K - kmin(s)], ai);
                });
                // (-1)^(N+m): m = 16 b + 4 q + (s % 4) has the parity of s % 4
                if constexpr ((N + s % 4) & 1) {
                    ar = -ar;
                    ai = -ai;
                }
                const int m = order(s);
                const bool on = m <= N;
                const float u = pw[2 * kMax * kMax + N * kMax + (on ? m : 0)];
                xr[s] = on ? ar * u : 0.0f;
                xi[s] = on ? ai * u : 0.0f;
            }
        });
        swap32<N, 2>(tab, lane, xr, xi);
        rot<N, false>(cb, sb, xr, xi);
        swap32<N, 2>(tab, lane, xr, xi);
        rot<N, false>(ca, sa, xr, xi);
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (kmin(s) <= N) {
                const int m = order(s);
                if (valid && m <= N) {
                    const double f = bs * static_cast<double>(pw[3 * kMax * kMax + N * kMax + m]);
                    D[N * (N + 1) + m] = cr[s] + static_cast<double>(xr[s]) * f;
                    D[(N + 1) * (N + 1) + m] = ci[s] + static_cast<double>(xi[s]) * f;
                }
            }
        });
        bs *= rinv * static_cast<double>(2 * (N + 1));
    });
}

// Persistent workgroups of kWaves waves: stage the FP32 blocks of degrees
// < P and the basis powers in LDS, then each wave takes groups of 16
// translations.
template <int P>
__global__ void __launch_bounds__(64 * kWaves)
    k_m2l_mfma_f32(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                   const sfmm_item *__restrict__ items, uint32_t n_items, int acc)
{
    constexpr int TAB = base32(P);
    __shared__ float tab[TAB];
    __shared__ float pw[kPow];
    for (int i = threadIdx.x; i < TAB; i += 64 * kWaves) tab[i] = kTab32.swap[i];
    for (int i = threadIdx.x; i < kPow; i += 64 * kWaves) pw[i] = kTab32.pw[i];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t groups = (n_items + 15u) / 16u;
    for (uint32_t g = blockIdx.x * kWaves + wave; g < groups; g += gridDim.x * kWaves) {
        // an opaque zero offset and lane per group keep the compiler from
        // hoisting table reads and lane-dependent indices out of the loop
        // into registers
        int base = 0, ln = lane;
        asm volatile("" : "+s"(base), "+v"(ln));
        m2l_group<P>(src_pool, dst_pool, items, n_items, g, tab + base, pw + base, ln, acc != 0);
    }
}

using fn_t = void (*)(const double *, double *, const sfmm_item *, uint32_t, int);

// Kernels of orders 1..16 (mfma_prec.hip), 17..20 (mfma_prec_mid.hip) and
// 21..24 (mfma_prec_hi.hip), compiled in three translation units; each
// returns nullptr outside its range.
fn_t kernel_lo(uint32_t P);
fn_t kernel_mid(uint32_t P);
fn_t kernel_hi(uint32_t P);

template <int FIRST, int... Ps>
constexpr auto make_fns(std::integer_sequence<int, Ps...>)
{
    struct T {
        fn_t f[sizeof...(Ps)];
    } t{{&k_m2l_mfma_f32<Ps + FIRST>...}};
    return t;
}

}  // namespace mprec
}  // namespace sfmm
