// Fused rotate / translate-along-z / rotate-back kernels for M2M, M2L and
// L2L on one MI355X (gfx950), FP64, one lane per translation.
//
// Algorithm (Dehnen's rotation-based O(P^3) scheme, as the reference
// implements it in translations.cpp:36-471):
//   forward, per input degree n:   scale, rotate about z by alpha, swap
//                                  x<->z, rotate by beta, swap x<->z
//   z-translation, per order m:    M2L Hankel sum over (N+K)!,
//                                  M2M/L2L Toeplitz sums over 1/(N-K)!
//   backward, per output degree N: swap, rotate by -beta, swap, rotate
//                                  by -alpha, scale, add into the target.
//
// Mapping: one lane owns one translation from start to end.  Degrees are
// unrolled at compile time, so every swap-matrix entry and factorial is an
// immediate operand (zero entries are skipped), and the rotated triangle of
// coefficients lives in registers; nothing but the source, the target and
// the item touches memory.  The swap contractions run on the FP64 VALU: on
// MI355X the FP64 MFMA peak equals the FP64 VALU peak on paper (78.6 TFLOP/s
// each; measured, FP64 FMA sustains 58 and v_mfma_f64_16x16x4f64 47)
// and the folded swap blocks are half zeros, which a VALU contraction skips
// and an MFMA tile cannot.
//
// Two variants per (op, PI, PO):
//   RT = false  every item of the launch has orders (PI, PO) exactly;
//   RT = true   PI = PO = PB is a bound and each item brings its own
//               p_in, p_out <= PB (mixed-order lists).  A missing source
//               degree is a zero, which leaves every output degree the
//               translation of order (p_in, p_out) defines unchanged
//               (the z-translation sums only over K < p_in); output degrees
//               >= p_out are not stored.  Degrees above the largest order
//               of the wave are skipped with wave-uniform branches, so a
//               wave of equal orders does the work of its own orders.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "../../include/solidfmm_gpu.h"
#include "tables.hpp"

namespace sfmm {
namespace lane {

constexpr int tri(int m, int n) { return n * (n + 1) / 2 + m; }

// e^{i alpha} and e^{i beta}.  The powers e^{i m theta} are formed inside
// each rotation by the reference's recurrence (microkernel_generic.cpp:
// 76-96), so no table of powers occupies registers.
struct Rot {
    double ca, sa, cb, sb;
};

// out[m] = sum_l A[m][l] in[l] for the folded real (KR) and imaginary (KI)
// blocks of degree N (microkernel_generic.cpp:142-214).
template <int N, int KR, int KI>
__device__ __forceinline__ void swap_stage(double (&re)[N + 1], double (&im)[N + 1])
{
    double ore[N + 1], oim[N + 1];
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int fr = first_nonzero<N, KR>(m);
        constexpr int fi = first_nonzero<N, KI>(m);
        double ar = 0.0, ai = 0.0;
        static_for<0, N + 1>([&](auto lc) {
            constexpr int l = decltype(lc)::value;
            constexpr double cr = folded<N, KR>(m, l);
            constexpr double ci = folded<N, KI>(m, l);
            if constexpr (cr != 0.0) {
                if constexpr (l == fr) ar = cr * re[l];
                else ar = __builtin_fma(cr, re[l], ar);
            }
            if constexpr (ci != 0.0) {
                if constexpr (l == fi) ai = ci * im[l];
                else ai = __builtin_fma(ci, im[l], ai);
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

// (re,im) <- scale * e^{+i m theta} (re,im) (FWD, rotscale forward=true)
// or scale * e^{-i m theta} (re,im); e^{i m theta} = (c1 + i s1)^m built by
// repeated products on the way.
template <int N, bool FWD, bool SCALE>
__device__ __forceinline__ void rot(double c1, double s1, double scale, double (&re)[N + 1], double (&im)[N + 1])
{
    if constexpr (SCALE) {
        re[0] *= scale;
        im[0] *= scale;
    }
    double c = c1, s = s1;
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if constexpr (m > 1) {
            const double c2 = c * c1 - s * s1;
            s = s * c1 + c * s1;
            c = c2;
        }
        const double a = re[m], b = im[m];
        double u, v;
        if constexpr (FWD) {
            u = c * a - s * b;
            v = s * a + c * b;
        } else {
            u = c * a + s * b;
            v = c * b - s * a;
        }
        re[m] = SCALE ? scale * u : u;
        im[m] = SCALE ? scale * v : v;
    });
}

// Issue the loads of degree N of a block into v[0..N] (16-byte pieces), or
// zeros when `on` is false.  Callers place a scheduling barrier after it, so
// the loads stay where they were issued and their latency overlaps the work
// that follows.
template <int N, int CAP>
__device__ __forceinline__ void fetch_degree(const double *__restrict__ blk, double2 (&v)[CAP], bool on)
{
    static_assert(N < CAP, "prefetch buffer too small");
    const double2 *q = reinterpret_cast<const double2 *>(blk + N * (N + 1));
    static_for<0, N + 1>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = on ? q[j] : double2{0.0, 0.0};
    });
}

// target degree N = old + (re, im), old prefetched by fetch_degree
// (buf2solid, microkernel_generic.cpp:361-380).  Nothing is stored when
// `on` is false.
template <int N, int CAP>
__device__ __forceinline__ void add_store_degree(double *__restrict__ blk, const double2 (&old)[CAP],
                                                 const double (&re)[N + 1], const double (&im)[N + 1], bool on)
{
    double2 *q = reinterpret_cast<double2 *>(blk + N * (N + 1));
    double a[2 * (N + 1)];
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        a[m] = re[m];
        a[N + 1 + m] = im[m];
    });
    if (on) {
        static_for<0, N + 1>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            double2 v = old[j];
            v.x += a[2 * j];
            v.y += a[2 * j + 1];
            q[j] = v;
        });
    }
}

// Degree N of a whole block held as 16-byte pieces b[0..P(P+1)/2).
template <int N, int CAP>
__device__ __forceinline__ void unpack_block_degree(const double2 (&b)[CAP], double (&re)[N + 1],
                                                    double (&im)[N + 1])
{
    static_assert(N * (N + 1) / 2 + N < CAP, "degree outside the block");
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        constexpr int kr = N * (N + 1) + m;          // Re C_N^m
        constexpr int ki = (N + 1) * (N + 1) + m;    // Im C_N^m
        re[m] = (kr & 1) ? b[kr / 2].y : b[kr / 2].x;
        im[m] = (ki & 1) ? b[ki / 2].y : b[ki / 2].x;
    });
}

__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// Largest value of v over the active lanes of the wave (0 <= v < 128),
// wave-uniform: one ballot per bit.
__device__ __forceinline__ uint32_t wave_max7(uint32_t v)
{
    uint32_t m = 0;
    for (int b = 6; b >= 0; --b) {
        const uint32_t cand = m | (1u << b);
        if (__ballot(v >= cand)) m = cand;
    }
    return m;
}

// Forward pipelines: M-type (M2M, M2L; translations.cpp:36-85) uses the
// folded blocks, L-type (L2L; translations.cpp:189-239) their transposes.
template <int N, bool LTYPE>
__device__ __forceinline__ void forward(const Rot &R, double scale, double (&re)[N + 1], double (&im)[N + 1])
{
    constexpr int KR = LTYPE ? 2 : 0;
    constexpr int KI = LTYPE ? 3 : 1;
    rot<N, true, true>(R.ca, R.sa, scale, re, im);
    swap_stage<N, KR, KI>(re, im);
    rot<N, true, false>(R.cb, R.sb, 1.0, re, im);
    swap_stage<N, KR, KI>(re, im);
}

// Backward pipelines: M-type (M2M; translations.cpp:87-136), L-type
// (M2L, L2L; translations.cpp:138-187).
template <int N, bool LTYPE>
__device__ __forceinline__ void backward(const Rot &R, double scale, double (&re)[N + 1], double (&im)[N + 1])
{
    constexpr int KR = LTYPE ? 2 : 0;
    constexpr int KI = LTYPE ? 3 : 1;
    swap_stage<N, KR, KI>(re, im);
    rot<N, false, false>(R.cb, R.sb, 1.0, re, im);
    swap_stage<N, KR, KI>(re, im);
    rot<N, false, true>(R.ca, R.sa, scale, re, im);
}

// Zeros read in place of a target that is not accumulated (acc == 0, or an
// output degree a per-item-order lane does not store): loads are never
// skipped, so no branch splits the load stream and the compiler can wait
// for each source piece separately.
inline __device__ double kZeroBlock[kMaxOrder * (kMaxOrder + 1)];

// Zero shift in M2M / L2L: target += source over the common degrees
// (translations.cpp:539-547, 596-604); target = source when !acc.
__device__ __forceinline__ void zero_shift_add(const double *__restrict__ S, double *__restrict__ D, uint32_t p,
                                               bool acc)
{
    const uint32_t len = p * (p + 1);
    for (uint32_t i = 0; i < len; ++i) D[i] = (acc ? D[i] : 0.0) + S[i];
}

// Segmented scratch mode: lanes whose items share a scratch block (a run of
// one target's items with equal orders, laid out by the plan) add their
// results by a fixed tree of shuffles and the run's first lane stores the
// sum, so a run of k items writes one block instead of k.
struct Segments {
    uint32_t same = 0;  // bit j: lane + 2^j belongs to this lane's run
    bool head = true;   // this lane stores its run's sum
};

__device__ __forceinline__ double shfl_down(double v, int d)
{
    return __shfl_down(v, static_cast<unsigned>(d), 64);
}

// Runs are identified by equal keys (the item's block offset; lanes without
// an item get a key of their own).  Call with all lanes of the wave active.
__device__ __forceinline__ Segments find_segments(uint64_t key)
{
    Segments g;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint32_t lo = static_cast<uint32_t>(key), hi = static_cast<uint32_t>(key >> 32);
    for (int j = 0; j < 6; ++j) {
        const int d = 1 << j;
        const uint32_t olo = __shfl_down(lo, static_cast<unsigned>(d), 64);
        const uint32_t ohi = __shfl_down(hi, static_cast<unsigned>(d), 64);
        if (lane + d < 64 && olo == lo && ohi == hi) g.same |= 1u << j;
    }
    const uint32_t plo = __shfl_up(lo, 1u, 64), phi = __shfl_up(hi, 1u, 64);
    g.head = lane == 0 || plo != lo || phi != hi;
    return g;
}

// v of the run's first lane becomes the sum over the run: after the step
// with offset d, lane l holds the sum over lanes l .. l + 2d - 1 of its run.
template <int N>
__device__ __forceinline__ void segment_sum(const Segments &g, double (&re)[N + 1], double (&im)[N + 1])
{
    static_for<0, 6>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const bool take = (g.same >> j) & 1u;
        static_for<0, N + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const double a = shfl_down(re[m], 1 << j);
            const double b = shfl_down(im[m], 1 << j);
            re[m] += take ? a : 0.0;
            im[m] += take ? b : 0.0;
        });
    });
}

// One lane, one translation (see the head of this file for RT).  acc = 1:
// target += T(source); acc = 0: target = T(source) over the degrees the
// translation defines (scratch blocks of a segmented reduction); acc = 2:
// as 0, with the lanes of a run adding into one block (Segments).
// Waves per SIMD the register allocation must allow.  At orders 7 and 8 the
// kernels need 256-316 registers unconstrained; held to the 256 that two
// waves per SIMD leave each lane (spilling up to 59 of them at order 8),
// they keep two waves in flight and run 19-26 % faster.  At orders 9 and
// 10 the spills (113-173 registers) cost more than the second wave gains.
template <int OP, int PI>
inline constexpr int kLaneWaves = (PI == 7 || PI == 8) ? 2 : 1;

// ipw: items per wave (1..64).  Wave w runs items [w ipw, (w+1) ipw) on its
// first ipw lanes; a launch sized below a whole number of rounds of resident
// waves spreads its items evenly over one more round instead of leaving
// that round's waves partly empty (launch_kernel).
template <int OP, int PI, int PO, bool RT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kLaneWaves<OP, PI>)))
    k_translate(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                const sfmm_item *__restrict__ items, uint32_t n_items, int acc, uint32_t ipw)
{
    static_assert(!RT || PI == PO, "runtime orders use a square bound");
    constexpr int PM = PI > PO ? PI : PO;
    constexpr int PE = PI < PO ? PI : PO;
    const uint32_t lane_id = threadIdx.x;
    const uint32_t t = lane_id < ipw ? blockIdx.x * ipw + lane_id : n_items;
    // runs (acc == 2) exist in the per-item-order kernels only; the plan
    // sends scratch launches with runs there, so the fixed-order kernels
    // carry no code for them
    Segments seg;
    if constexpr (RT) {
        if (acc == 2) {  // before any lane leaves
            uint64_t key = ~uint64_t(threadIdx.x & 63);
            if (t < n_items) key = items[t].dst;
            seg = find_segments(key);
        }
    }
    if (t >= n_items) return;

    const sfmm_item it = items[t];
    const double *__restrict__ S = static_cast<const double *>(__builtin_assume_aligned(src_pool + it.src, 16));
    double *__restrict__ D = static_cast<double *>(__builtin_assume_aligned(dst_pool + it.dst, 16));
    double x = it.x, y = it.y, z = it.z;
    // This lane's orders; for L2L the outputs stop at min(p_in, p_out)
    // (translations.cpp:412).
    const uint32_t pin = RT ? it.p_in : uint32_t(PI);
    const uint32_t pout = RT ? it.p_out : uint32_t(PO);
    const uint32_t pend = pin < pout ? pin : pout;

    if constexpr (OP != SFMM_M2L) {
        if (x * x + y * y + z * z == 0.0) {
            zero_shift_add(S, D, RT ? pend : uint32_t(PE), acc == 1);
            return;
        }
    }
    if constexpr (OP == SFMM_M2M) {  // translations.cpp:331-333
        x = -x;
        y = -y;
        z = -z;
    }
    // Wave-uniform bounds of the degrees worth computing.
    const uint32_t pin_w = RT ? wave_max7(pin) : uint32_t(PI);
    const uint32_t pout_w = RT ? wave_max7(OP == SFMM_L2L ? pend : pout) : uint32_t(PO);
    auto out_on = [&](int N) { return !RT || uint32_t(N) < (OP == SFMM_L2L ? pend : pout); };
    // where degree N's old target values come from: the target, or zeros
    const double *__restrict__ Dr = acc == 1 ? D : kZeroBlock;
    // store degree N: in segmented mode the run's sum, by its first lane
    auto finish_degree = [&](auto Nc, const double2(&cur)[PO], double(&re)[decltype(Nc)::value + 1],
                             double(&im)[decltype(Nc)::value + 1], bool on) {
        constexpr int N = decltype(Nc)::value;
        if constexpr (RT) {
            if (acc == 2) {
                segment_sum<N>(seg, re, im);
                on = on && seg.head;
            }
        }
        add_store_degree<N>(D, cur, re, im, on);
    };
    auto read_from = [&](int N) { return out_on(N) ? Dr : kZeroBlock; };

    // Euler data (microkernel_generic.cpp:61-74).
    const double rho2 = x * x + y * y;
    const double rho = sqrt(rho2);
    const double r = sqrt(rho2 + z * z);
    const double rinv = 1.0 / r;
    double ca1 = 1.0, sa1 = 0.0;
    if (rho != 0.0) {
        const double q = 1.0 / rho;
        ca1 = y * q;
        sa1 = x * q;
    }
    const Rot R{ca1, sa1, z * rinv, -rho * rinv};

    // Whole source block in flight at once.  Its registers are handed over
    // degree by degree to the rotated triangle, so this costs no peak.
    double2 sv[tri(0, PI)];
    {
        const double2 *q = reinterpret_cast<const double2 *>(S);
        const uint32_t have = pin * (pin + 1) / 2;  // 16-byte pieces of degrees < p_in
        static_for<0, tri(0, PI)>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            sv[j] = (!RT || uint32_t(j) < have) ? q[j] : double2{0.0, 0.0};
        });
    }
    // Target degrees are prefetched kAhead degrees ahead of their use, into
    // a ring of buffers (slot = degree mod kAhead); the scheduling fences
    // keep those loads where they are issued.
    // One degree ahead: at P=10 M2L, prefetching 2, 3 or all 10 target
    // degrees measured 1.5 %, 10 % and 28 % slower (register pressure).
    constexpr int kAhead = 1;
    double2 tg[kAhead][PO];
    auto take = [&](auto Nc, double2(&cur)[PO]) {
        constexpr int N = decltype(Nc)::value;
        static_for<0, N + 1>([&](auto jc) { cur[decltype(jc)::value] = tg[N % kAhead][decltype(jc)::value]; });
    };
    auto prefetch = [&](auto Nc) {
        constexpr int N = decltype(Nc)::value;
        if constexpr (N >= 0 && N < PO) fetch_degree<N>(read_from(N), tg[N % kAhead], true);
    };

    // Rotated triangle: T[tri(m,K)] holds order m of degree K.
    double tre[tri(0, PI)], tim[tri(0, PI)];
    auto keep = [&](auto nc, const auto &re, const auto &im) {
        constexpr int n = decltype(nc)::value;
        static_for<0, n + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            tre[tri(m, n)] = re[m];
            tim[tri(m, n)] = im[m];
        });
    };
    auto clear = [&](auto nc) {
        constexpr int n = decltype(nc)::value;
        static_for<0, n + 1>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            tre[tri(m, n)] = 0.0;
            tim[tri(m, n)] = 0.0;
        });
    };

    if constexpr (OP == SFMM_M2L) {
        // Forward, all input degrees; scale r^-(n+1) (translations.cpp:45).
        double s = rinv;
        static_for<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            if constexpr (n == PI - 1) {
                static_for<0, kAhead>([&](auto dc) { prefetch(ic<PO - 1 - decltype(dc)::value>{}); });
                sched_fence();
            }
            if (!RT || uint32_t(n) < pin_w) {
                double re[n + 1], im[n + 1];
                unpack_block_degree<n>(sv, re, im);
                forward<n, false>(R, s, re, im);
                keep(nc, re, im);
            } else {
                clear(nc);
            }
            sched_fence();
            s *= rinv;
        });
        double rp[PO];  // r^-N, the backward scale (translations.cpp:180)
        rp[0] = 1.0;
        static_for<1, PO>([&](auto Nc) {
            constexpr int N = decltype(Nc)::value;
            rp[N] = rp[N - 1] * rinv;
        });
        // z-translation and backward, highest degree first so rotated
        // columns die early.  L_N^m = (-1)^(N+m) sum_K (N+K)! T_K^m
        // (translations.cpp:283-296, zm2l microkernel_generic.cpp:216-255).
        static_for_down<0, PO>([&](auto Nc) {
            constexpr int N = decltype(Nc)::value;
            double2 cur[PO];
            take(Nc, cur);
            prefetch(ic<N - kAhead>{});
            sched_fence();
            if (!RT || uint32_t(N) < pout_w) {
                double re[N + 1], im[N + 1];
                static_for<0, N + 1>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    double ar = 0.0, ai = 0.0;
                    if constexpr (m < PI) {
                        static_for<m, PI>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value;
                            constexpr double f = kFac.fac[N + K];
                            if constexpr (K == m) {
                                ar = f * tre[tri(m, K)];
                                ai = f * tim[tri(m, K)];
                            } else {
                                ar = __builtin_fma(f, tre[tri(m, K)], ar);
                                ai = __builtin_fma(f, tim[tri(m, K)], ai);
                            }
                        });
                        if constexpr ((N + m) & 1) {
                            ar = -ar;
                            ai = -ai;
                        }
                    }
                    re[m] = ar;
                    im[m] = ai;
                });
                backward<N, true>(R, rp[N], re, im);
                finish_degree(ic<N>{}, cur, re, im, out_on(N));
            }
        });
    } else if constexpr (OP == SFMM_M2M) {
        // Ascending degrees; output degree n needs input degrees K <= n
        // only, so each is finished as soon as its inputs exist.
        // M_N^m = sum_{K<=N} r^{N-K}/(N-K)! T_K^m (translations.cpp:363-377).
        double s = rinv;   // forward scale r^-(n+1)
        double sb = r;     // backward scale r^(N+1) (translations.cpp:129)
        static_for<0, kAhead>([&](auto dc) { prefetch(ic<decltype(dc)::value>{}); });
        sched_fence();
        static_for<0, PM>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            double2 cur[PO];
            if constexpr (n < PO) {
                take(nc, cur);
                prefetch(ic<n + kAhead>{});
                sched_fence();
            }
            if constexpr (n < PI) {
                if (!RT || uint32_t(n) < pin_w) {
                    double re[n + 1], im[n + 1];
                    unpack_block_degree<n>(sv, re, im);
                    forward<n, false>(R, s, re, im);
                    keep(nc, re, im);
                } else {
                    clear(nc);
                }
                sched_fence();
                s *= rinv;
            }
            if constexpr (n < PO) {
                constexpr int N = n;
                if (!RT || uint32_t(N) < pout_w) {
                    double re[N + 1], im[N + 1];
                    static_for<0, N + 1>([&](auto mc) {
                        constexpr int m = decltype(mc)::value;
                        double ar = 0.0, ai = 0.0;
                        if constexpr (m < PI) {
                            constexpr int KE = (N + 1 < PI) ? N + 1 : PI;
                            static_for<m, KE>([&](auto Kc) {
                                constexpr int K = decltype(Kc)::value;
                                constexpr double f = kFac.inv[N - K];
                                if constexpr (K == m) {
                                    ar = f * tre[tri(m, K)];
                                    ai = f * tim[tri(m, K)];
                                } else {
                                    ar = __builtin_fma(f, tre[tri(m, K)], ar);
                                    ai = __builtin_fma(f, tim[tri(m, K)], ai);
                                }
                            });
                        }
                        re[m] = ar;
                        im[m] = ai;
                    });
                    backward<N, false>(R, sb, re, im);
                    finish_degree(ic<N>{}, cur, re, im, out_on(N));
                }
                sb *= r;
            }
        });
    } else {
        // L2L, descending degrees; output degree n needs input degrees
        // K >= n only.  L_N^m = sum_{K>=N} r^{K-N}/(K-N)! T_K^m
        // (translations.cpp:440-451).  Forward scale r^n (:198), backward
        // r^-n (:180).
        double fs[PI];
        fs[0] = 1.0;
        static_for<1, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            fs[n] = fs[n - 1] * r;
        });
        double bs[PE > 0 ? PE : 1];
        bs[0] = 1.0;
        static_for<1, PE>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            bs[n] = bs[n - 1] * rinv;
        });
        static_for<0, kAhead>([&](auto dc) { prefetch(ic<PE - 1 - decltype(dc)::value>{}); });
        sched_fence();
        static_for_down<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            if (!RT || uint32_t(n) < pin_w) {
                double re[n + 1], im[n + 1];
                unpack_block_degree<n>(sv, re, im);
                forward<n, true>(R, fs[n], re, im);
                keep(nc, re, im);
            } else {
                clear(nc);
            }
            sched_fence();
            if constexpr (n < PE) {
                constexpr int N = n;
                double2 cur[PO];
                take(ic<N>{}, cur);
                prefetch(ic<N - kAhead>{});
                sched_fence();
                if (!RT || uint32_t(N) < pout_w) {
                    double re[N + 1], im[N + 1];
                    static_for<0, N + 1>([&](auto mc) {
                        constexpr int m = decltype(mc)::value;
                        double ar = 0.0, ai = 0.0;
                        static_for<N, PI>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value;
                            constexpr double f = kFac.inv[K - N];
                            if constexpr (K == N) {
                                ar = tre[tri(m, K)];
                                ai = tim[tri(m, K)];
                            } else {
                                ar = __builtin_fma(f, tre[tri(m, K)], ar);
                                ai = __builtin_fma(f, tim[tri(m, K)], ai);
                            }
                        });
                        re[m] = ar;
                        im[m] = ai;
                    });
                    backward<N, true>(R, bs[N], re, im);
                    finish_degree(ic<N>{}, cur, re, im, out_on(N));
                }
            }
        });
    }
}

using kernel_fn = void (*)(const double *, double *, const sfmm_item *, uint32_t, int, uint32_t);

// Square orders 1..kSquare, all three operators.
inline constexpr int kSquare = 12;

struct Table {
    kernel_fn f[3][kSquare];
};

template <bool RT, int OP, int... Ps>
constexpr void fill(kernel_fn *tab, std::integer_sequence<int, Ps...>)
{
    ((tab[Ps] = &k_translate<OP, Ps + 1, Ps + 1, RT>), ...);
}

template <bool RT>
Table make_table()
{
    Table t{};
    fill<RT, SFMM_M2M>(t.f[0], std::make_integer_sequence<int, kSquare>{});
    fill<RT, SFMM_M2L>(t.f[1], std::make_integer_sequence<int, kSquare>{});
    fill<RT, SFMM_L2L>(t.f[2], std::make_integer_sequence<int, kSquare>{});
    return t;
}

// Tables of the two variants (lane_fixed.hip, lane_rt.hip).
const Table &fixed_table();
const Table &rt_table();

}  // namespace lane
}  // namespace sfmm
