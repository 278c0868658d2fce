// FP64 translation kernels whose x<->z swap contractions run on the matrix
// cores (v_mfma_f64_16x16x4f64).  Used for the orders the one-lane-per-
// translation kernels of translate.hip cannot hold in registers.
//
// Mapping.  A wave owns 16 translations; lane l = t + 16 q handles
// translation t and the orders m = q + 4 s ("slot" s) of every degree.  That
// is exactly how the MFMA spreads its 16x16 result over the wave (element s
// of lane (q,t) is row q + 4 s, column t), and it is also how the next MFMA
// wants its B operand (rows 4c..4c+3 of K-chunk c sit in element c of the
// four quarter-waves).  So a swap is
//     D[m][t] = sum_c  A_c[m][4c..4c+3] * x[4c..4c+3][t]
// with x taken straight from the previous result registers: no shuffles
// between the rotation, swap, rotation, swap steps.
//
// The A operands (the folded swap blocks, operator_data.cpp:47-147) are
// constant; each workgroup copies the dense table of mfma_tables.hpp for
// its degrees into LDS once and then loops over groups of 16 translations
// (persistent workgroups), so every operand read is an LDS read.
//
// The rotations e^{+-i m theta}, the z-translation sums and the scaling run
// on the FP64 VALU on the same registers; each lane keeps the rotated
// triangle only for its own orders m = q + 4s, so the z-translation of an
// order m is lane-local.  Algorithm: Dehnen's rotation scheme as in
// translations.cpp:36-471 (see translate.hip for the per-step citations).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "tables.hpp"
#include "translate.hpp"

namespace sfmm {
namespace mfma {

typedef double dbl4 __attribute__((ext_vector_type(4)));

inline constexpr int kWaves = 4;  // waves per workgroup

// Triangle storage per lane: slot s holds degrees K in [4s, PI).
template <int PI>
constexpr int slot_off(int s)
{
    int o = 0;
    for (int j = 0; j < s; ++j) o += PI - 4 * j;
    return o;
}
template <int PI>
constexpr int tri_size()
{
    return slot_off<PI>((PI + 3) / 4);
}

template <int P>
constexpr int nslots()
{
    return (P + 3) / 4;
}

// Orders whose slot 0 of the rotated triangle (PI reals and PI imaginary
// parts per lane) lives in LDS instead of registers: order 17, where the
// register file would otherwise overflow into scratch memory.  From 18 on
// the swaps recompute their indices (opaque_lane), which frees enough
// registers that the slot stays in registers without spills; that measured
// 3-6 % faster than the LDS slot at orders 18, 20 and 22.
template <int PI>
constexpr bool lds_slot0()
{
    return PI == 17;
}
template <int PI>
constexpr int lds_slot0_size()  // doubles per wave
{
    return lds_slot0<PI>() ? 2 * PI * 64 : 0;
}

__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Orders whose swaps recompute their lane-dependent indices and masks at
// every swap (OPQ).  Otherwise the compiler computes them once per kernel
// and holds them in registers for the whole translation: about 100
// registers at order 16.  Recomputing costs VALU instructions, so it is
// used where the registers it frees buy something: two waves per SIMD (M2L
// up to order 12: 8.96e8 -> 9.81e8/s at 12; M2M and L2L up to 10) or no
// spills (from order 18: the per-item-order M2L kernels spilled 124-624
// bytes per lane at 19-22 and spill none).  From 13 to 17 the kernels stay
// at one wave per SIMD either way and run 2-8 % faster without it.
template <int OP, int PI>
constexpr bool opaque_lane()
{
    return PI <= (OP == SFMM_M2L ? 12 : 10) || PI >= 18;
}

// One swap step for degree N on the lane's slot values (xr, xi), with the
// folded real and imaginary blocks (TR = false) or their transposes
// (TR = true, via D^-1 S^T D).  tab points at the LDS copy of the table.
template <int N, bool TR, bool OPQ, int NSL>
__device__ __forceinline__ void swap_mfma(const double *tab, int lane, double (&xr)[NSL], double (&xi)[NSL])
{
    constexpr int NB = (N + 16) / 16;  // row blocks of 16 orders
    constexpr int NC = (N + 4) / 4;    // K-chunks of 4 orders
    constexpr int ST = dense_stride(N);
    const double *sr = tab + dense_base(N);
    const double *si = sr + dense_part(N);
    int ln = lane;
    if constexpr (OPQ) asm volatile("" : "+v"(ln));  // see opaque_lane
    const int r = ln & 15;  // operand row within the block (output order)
    const int k = ln >> 4;  // operand column within the chunk (input order); also q

    if constexpr (TR) {
        // inputs of order l = 4c + q >= 1 doubled
        static_for<0, NC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            const double f = (c > 0 || k > 0) ? 2.0 : 1.0;  // a select, not a branch
            xr[c] *= f;
            xi[c] *= f;
        });
    }
    dbl4 dr[NB], di[NB];
    static_for<0, NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        dr[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        di[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        const int m = 16 * b + r;
        static_for<0, NC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            const int l = 4 * c + k;
            const bool in = (m <= N) && (l <= N);
            const int idx = TR ? (l * ST + m) : (m * ST + l);
            const int safe = in ? idx : 0;
            // always read (a valid entry), then mask by a factor: a select
            // here would be turned into a branch around the LDS read
            const double fin = in ? 1.0 : 0.0;
            const double ar = sr[safe] * fin;
            const double ai = si[safe] * fin;
            dr[b] = mfma(ar, xr[c], dr[b]);
            di[b] = mfma(ai, xi[c], di[b]);
        });
    });
    static_for<0, NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        static_for<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if constexpr (4 * b + i < NSL) {
                xr[4 * b + i] = dr[b][i];
                xi[4 * b + i] = di[b][i];
            }
        });
    });
    if constexpr (TR) {
        // outputs of order m = q + 4s + 16b >= 1 halved
        static_for<0, 4 * NB>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (s < NSL) {
                const double f = (s > 0 || k > 0) ? 0.5 : 1.0;
                xr[s] *= f;
                xi[s] *= f;
            }
        });
    }
}

// Per-lane rotation factors e^{i m alpha}, e^{i m beta} for m = q + 4s.
template <int NSL>
struct LaneRot {
    double ca[NSL], sa[NSL], cb[NSL], sb[NSL];
};

template <int N, bool BETA, bool FWD, int NSL>
__device__ __forceinline__ void rot(const LaneRot<NSL> &R, double scale, double (&xr)[NSL], double (&xi)[NSL])
{
    constexpr int NS = (N + 4) / 4;
    static_for<0, NS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const double c = BETA ? R.cb[s] : R.ca[s];
        const double sn = BETA ? R.sb[s] : R.sa[s];
        const double a = xr[s], b = xi[s];
        if constexpr (FWD) {
            xr[s] = scale * (c * a - sn * b);
            xi[s] = scale * (sn * a + c * b);
        } else {
            xr[s] = scale * (c * a + sn * b);
            xi[s] = scale * (c * b - sn * a);
        }
    });
}

template <int N, bool TR, bool OPQ, int NSL>
__device__ __forceinline__ void transform_fwd(const double *tab, int lane, const LaneRot<NSL> &R, double scale,
                                              double (&xr)[NSL], double (&xi)[NSL])
{
    rot<N, false, true>(R, scale, xr, xi);
    swap_mfma<N, TR, OPQ>(tab, lane, xr, xi);
    rot<N, true, true>(R, 1.0, xr, xi);
    swap_mfma<N, TR, OPQ>(tab, lane, xr, xi);
}

template <int N, bool TR, bool OPQ, int NSL>
__device__ __forceinline__ void transform_bwd(const double *tab, int lane, const LaneRot<NSL> &R, double scale,
                                              double (&xr)[NSL], double (&xi)[NSL])
{
    swap_mfma<N, TR, OPQ>(tab, lane, xr, xi);
    rot<N, true, false>(R, 1.0, xr, xi);
    swap_mfma<N, TR, OPQ>(tab, lane, xr, xi);
    rot<N, false, false>(R, scale, xr, xi);
}

// Read in place of a coefficient a lane does not load.
inline __device__ double kZeros[2];

// Largest value of v over the wave (0 <= v < 128), wave-uniform: one
// ballot per bit.
__device__ __forceinline__ uint32_t wave_max7(uint32_t v)
{
    uint32_t m = 0;
    for (int b = 6; b >= 0; --b) {
        const uint32_t cand = m | (1u << b);
        if (__ballot(v >= cand)) m = cand;
    }
    return m;
}

// Translations of one group of 16 (items [16g, 16g+16) of the launch).
// RT = false: every item has orders (PI, PO).  RT = true: PI = PO is a
// bound and each item brings its own p_in, p_out (mixed-order lists); a
// missing source degree is a zero, output degrees >= p_out (L2L: >=
// min(p_in, p_out), translations.cpp:412) are not stored, and degrees above
// the group's largest orders are skipped with wave-uniform branches.
template <int OP, int PI, int PO, bool RT>
__device__ __forceinline__ void translate_group(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                                                const sfmm_item *__restrict__ items, uint32_t n_items, uint32_t g,
                                                const double *tab, double *tl, int lane, int mode)
{
    // tl: this wave's LDS copy of slot 0 of the triangle when
    // lds_slot0<PI>(): Re of degree K at tl[64 K + lane], Im at
    // tl[64 (PI + K) + lane]; each lane reads back only what it wrote.
    // mode 1: target += T(source); 0: scratch block = T(source); 2: as 0,
    // with the items of a run (consecutive translations of the group with
    // one scratch block) added across lanes and stored by the run's first
    const bool acc = mode == 1;
    constexpr int PM = PI > PO ? PI : PO;
    constexpr int PE = PI < PO ? PI : PO;
    constexpr int NSL = 4 * ((PM + 15) / 16);  // slots incl. padding to whole row blocks
    const int q = lane >> 4;
    const int t = lane & 15;
    const uint32_t idx = g * 16u + static_cast<uint32_t>(t);
    const bool valid = idx < n_items;

    // Every lane stays active to the end: the MFMAs need the whole wave.
    sfmm_item it{};
    if (valid) it = items[idx];
    // runs within the group: bit j of `same` says translation t + 2^j is in
    // t's run; the run's first translation stores the sum
    uint32_t same = 0;
    bool head = true;
    if (mode == 2) {
        const uint64_t key = valid ? it.dst : ~uint64_t(lane);
        const uint32_t lo = uint32_t(key), hi = uint32_t(key >> 32);
        static_for<0, 4>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const uint32_t olo = __shfl_down(lo, 1u << j, 64), ohi = __shfl_down(hi, 1u << j, 64);
            if (t + (1 << j) < 16 && olo == lo && ohi == hi) same |= 1u << j;
        });
        const uint32_t plo = __shfl_up(lo, 1u, 64), phi = __shfl_up(hi, 1u, 64);
        head = t == 0 || plo != lo || phi != hi;
    }
    const double *__restrict__ S = src_pool + it.src;
    double *__restrict__ D = dst_pool + it.dst;
    double x = it.x, y = it.y, z = it.z;

    bool zero_shift = false;
    if constexpr (OP != SFMM_M2L) zero_shift = valid && (x * x + y * y + z * z == 0.0);
    if constexpr (OP == SFMM_M2M) {  // translations.cpp:331-333
        x = -x;
        y = -y;
        z = -z;
    }
    if (!valid || zero_shift) {  // harmless geometry for idle lanes
        x = 0.0;
        y = 0.0;
        z = 1.0;
    }
    const bool active = valid && !zero_shift;
    const uint32_t pin = RT ? it.p_in : uint32_t(PI);
    const uint32_t pout = RT ? it.p_out : uint32_t(PO);
    const uint32_t pend = pin < pout ? pin : pout;
    const uint32_t pin_w = RT ? wave_max7(active ? pin : 0u) : uint32_t(PI);
    const uint32_t pout_w = RT ? wave_max7(active ? (OP == SFMM_L2L ? pend : pout) : 0u) : uint32_t(PO);
    // does this lane store output degree N?
    auto out_on = [&](int N) { return !RT || uint32_t(N) < (OP == SFMM_L2L ? pend : pout); };

    // Euler data (microkernel_generic.cpp:61-74).
    const double rho2 = x * x + y * y;
    const double rho = sqrt(rho2);
    const double r = sqrt(rho2 + z * z);
    const double rinv = 1.0 / r;
    double ca1 = 1.0, sa1 = 0.0;
    if (rho != 0.0) {
        const double qq = 1.0 / rho;
        ca1 = y * qq;
        sa1 = x * qq;
    }
    const double cb1 = z * rinv, sb1 = -rho * rinv;

    // e^{i k theta} for k = 0..PM-1 by repeated products (as the reference
    // builds them, microkernel_generic.cpp:76-96); the lane keeps k = q+4s.
    LaneRot<NSL> R{};
    {
        double c = 1.0, s = 0.0, cb = 1.0, sb = 0.0;
        static_for<0, 4 * NSL>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int slot = k / 4;
            const bool mine = (k & 3) == q;  // selects, not branches
            R.ca[slot] = mine ? c : R.ca[slot];
            R.sa[slot] = mine ? s : R.sa[slot];
            R.cb[slot] = mine ? cb : R.cb[slot];
            R.sb[slot] = mine ? sb : R.sb[slot];
            if constexpr (k + 1 < 4 * NSL) {
                if constexpr (k == 0) {
                    c = ca1;
                    s = sa1;
                    cb = cb1;
                    sb = sb1;
                } else {
                    const double c2 = c * ca1 - s * sa1;
                    s = s * ca1 + c * sa1;
                    c = c2;
                    const double cb2 = cb * cb1 - sb * sb1;
                    sb = sb * cb1 + cb * sb1;
                    cb = cb2;
                }
            }
        });
    }

    // Rotated triangle, this lane's orders only (slot 0 in LDS when
    // lds_slot0<PI>(); its register copies are then never touched).
    double tr[tri_size<PI>()], ti[tri_size<PI>()];
    auto put_tri = [&](auto sc, auto Kc, double re, double im) {
        constexpr int s = decltype(sc)::value;
        constexpr int K = decltype(Kc)::value;
        if constexpr (s == 0 && lds_slot0<PI>()) {
            tl[64 * K + lane] = re;
            tl[64 * (PI + K) + lane] = im;
        } else {
            tr[slot_off<PI>(s) + K - 4 * s] = re;
            ti[slot_off<PI>(s) + K - 4 * s] = im;
        }
    };
    auto tri_re = [&](auto sc, auto Kc) -> double {
        constexpr int s = decltype(sc)::value;
        constexpr int K = decltype(Kc)::value;
        if constexpr (s == 0 && lds_slot0<PI>()) return tl[64 * K + lane];
        else return tr[slot_off<PI>(s) + K - 4 * s];
    };
    auto tri_im = [&](auto sc, auto Kc) -> double {
        constexpr int s = decltype(sc)::value;
        constexpr int K = decltype(Kc)::value;
        if constexpr (s == 0 && lds_slot0<PI>()) return tl[64 * (PI + K) + lane];
        else return ti[slot_off<PI>(s) + K - 4 * s];
    };

    auto load_degree = [&](auto nc, double(&xr)[NSL], double(&xi)[NSL]) {
        constexpr int n = decltype(nc)::value;
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const int m = q + 4 * s;
            xr[s] = 0.0;
            xi[s] = 0.0;
            if constexpr (4 * s <= n) {
                // masked lanes read zeros instead of skipping the loads, so no
                // branch splits the load stream
                const bool ok = active && m <= n && (!RT || uint32_t(n) < pin);
                const double *p = ok ? S + (n * (n + 1) + m) : kZeros;
                const int im = ok ? n + 1 : 0;  // Im C_n^m follows Re C_n^m by n+1
                xr[s] = p[0];
                xi[s] = p[im];
            }
        });
    };
    auto keep_degree = [&](auto nc, const double(&xr)[NSL], const double(&xi)[NSL]) {
        constexpr int n = decltype(nc)::value;
        static_for<0, nslots<PI>()>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (4 * s <= n) put_tri(sc, nc, xr[s], xi[s]);
        });
    };
    auto clear_degree = [&](auto nc) {
        constexpr int n = decltype(nc)::value;
        static_for<0, nslots<PI>()>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (4 * s <= n) put_tri(sc, nc, 0.0, 0.0);
        });
    };
    auto fetch_old = [&](auto Nc, double(&orr)[NSL], double(&oi)[NSL]) {
        constexpr int N = decltype(Nc)::value;
        const bool on = out_on(N);
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            orr[s] = 0.0;
            oi[s] = 0.0;
            if constexpr (4 * s <= N) {
                const int m = q + 4 * s;
                const bool ok = active && m <= N && on && acc;
                const double *p = ok ? D + (N * (N + 1) + m) : kZeros;
                const int im = ok ? N + 1 : 0;
                orr[s] = p[0];
                oi[s] = p[im];
            }
        });
    };
    auto store_sum = [&](auto Nc, const double(&orr)[NSL], const double(&oi)[NSL], double(&xr)[NSL],
                         double(&xi)[NSL]) {
        constexpr int N = decltype(Nc)::value;
        bool on = out_on(N);
        if (mode == 2) {  // fixed tree over the run's translations (lanes t + 2^j, same q)
            static_for<0, 4>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const bool take = (same >> j) & 1u;
                static_for<0, NSL>([&](auto sc) {
                    constexpr int s = decltype(sc)::value;
                    if constexpr (4 * s <= N) {
                        const double a = __shfl_down(xr[s], 1u << j, 64);
                        const double b = __shfl_down(xi[s], 1u << j, 64);
                        xr[s] += take ? a : 0.0;
                        xi[s] += take ? b : 0.0;
                    }
                });
            });
            on = on && head;
        }
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (4 * s <= N) {
                const int m = q + 4 * s;
                if (active && m <= N && on) {
                    D[N * (N + 1) + m] = orr[s] + xr[s];
                    D[(N + 1) * (N + 1) + m] = oi[s] + xi[s];
                }
            }
        });
    };
    // Source and target values are fetched one degree ahead of their use;
    // the scheduling fences keep each fetch where it is issued.
    double pr[NSL], pim[NSL];  // prefetched source degree
    double orr[NSL], oi[NSL];  // prefetched target degree
    auto take = [](double(&to)[NSL], const double(&from)[NSL]) {
        static_for<0, NSL>([&](auto sc) { to[decltype(sc)::value] = from[decltype(sc)::value]; });
    };

    if constexpr (OP == SFMM_M2L) {
        // Output degrees run ascending in the per-item-order kernels, with the
        // backward scale r^-N as one running product; a table of PO powers
        // would hold PO registers through the loop (order 17 spilled).  The
        // fixed-order kernels keep descending degrees and the table: the
        // compiler schedules their straight-line code with fewer registers
        // that way (ascending spilled from order 14 on).
        constexpr bool kUp = RT;
        double sc = rinv;  // r^-(n+1) (translations.cpp:45)
        load_degree(ic<0>{}, pr, pim);
        static_for<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            double xr[NSL], xi[NSL];
            take(xr, pr);
            take(xi, pim);
            if constexpr (n + 1 < PI) load_degree(ic<n + 1>{}, pr, pim);
            else fetch_old(ic<kUp ? 0 : PO - 1>{}, orr, oi);
            sched_fence();
            if (!RT || uint32_t(n) < pin_w) {
                transform_fwd<n, false, opaque_lane<OP, PI>()>(tab, lane, R, sc, xr, xi);
                keep_degree(nc, xr, xi);
            } else {
                clear_degree(nc);
            }
            sc *= rinv;
        });
        auto out_degree = [&](auto Nc, double scale) {
            constexpr int N = decltype(Nc)::value;
            double cr[NSL], ci[NSL];
            take(cr, orr);
            take(ci, oi);
            if constexpr (kUp && N + 1 < PO) fetch_old(ic<N + 1>{}, orr, oi);
            if constexpr (!kUp && N > 0) fetch_old(ic<N - 1>{}, orr, oi);
            sched_fence();
            if (!RT || uint32_t(N) < pout_w) {
                double xr[NSL], xi[NSL];
                const bool odd = ((N + q) & 1) != 0;  // (-1)^(N+m), m = q + 4s
                static_for<0, NSL>([&](auto sc) {
                    constexpr int s = decltype(sc)::value;
                    double ar = 0.0, ai = 0.0;
                    if constexpr (4 * s <= N && 4 * s < PI) {
                        static_for<4 * s, PI>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value;
                            constexpr double f = kFac.fac[N + K];
                            ar = __builtin_fma(f, tri_re(sc, Kc), ar);
                            ai = __builtin_fma(f, tri_im(sc, Kc), ai);
                        });
                        ar = odd ? -ar : ar;
                        ai = odd ? -ai : ai;
                        ar = (q + 4 * s > N) ? 0.0 : ar;
                        ai = (q + 4 * s > N) ? 0.0 : ai;
                    }
                    xr[s] = ar;
                    xi[s] = ai;
                });
                transform_bwd<N, true, opaque_lane<OP, PI>()>(tab, lane, R, scale, xr, xi);
                store_sum(Nc, cr, ci, xr, xi);
            }
        };
        // backward scale r^-N: the same running products either way
        if constexpr (kUp) {
            double rpN = 1.0;
            static_for<0, PO>([&](auto Nc) {
                out_degree(Nc, rpN);
                rpN *= rinv;
            });
        } else {
            double rp[PO];
            rp[0] = 1.0;
            static_for<1, PO>([&](auto Nc) {
                constexpr int N = decltype(Nc)::value;
                rp[N] = rp[N - 1] * rinv;
            });
            static_for_down<0, PO>([&](auto Nc) { out_degree(Nc, rp[decltype(Nc)::value]); });
        }
    } else if constexpr (OP == SFMM_M2M) {
        // Stage s runs the forward transform of source degree s beside the
        // z-translation and backward transform of output degree s - 1
        // (which needs source degrees K <= s - 1 only): two independent
        // chains of MFMAs per stage.
        double sc = rinv;  // forward r^-(n+1)
        double sb = r;     // backward r^(N+1) (translations.cpp:129)
        load_degree(ic<0>{}, pr, pim);
        static_for<0, PM + 1>([&](auto stc) {
            constexpr int st = decltype(stc)::value;
            constexpr int n = st;      // forward degree
            constexpr int N = st - 1;  // output degree
            double xr[NSL], xi[NSL], cr[NSL], ci[NSL];
            if constexpr (n < PI) {
                take(xr, pr);
                take(xi, pim);
                if constexpr (n + 1 < PI) load_degree(ic<n + 1>{}, pr, pim);
            }
            if constexpr (N >= 0 && N < PO) {
                take(cr, orr);
                take(ci, oi);
            }
            if constexpr (N + 1 < PO) fetch_old(ic<N + 1>{}, orr, oi);
            sched_fence();
            if constexpr (n < PI) {
                if (!RT || uint32_t(n) < pin_w) {
                    transform_fwd<n, false, opaque_lane<OP, PI>()>(tab, lane, R, sc, xr, xi);
                    keep_degree(ic<n>{}, xr, xi);
                } else {
                    clear_degree(ic<n>{});
                }
                sc *= rinv;
            }
            if constexpr (N >= 0 && N < PO) {
                if (!RT || uint32_t(N) < pout_w) {
                    double yr[NSL], yi[NSL];
                    static_for<0, NSL>([&](auto scc) {
                        constexpr int s = decltype(scc)::value;
                        double ar = 0.0, ai = 0.0;
                        if constexpr (4 * s <= N && 4 * s < PI) {
                            constexpr int KE = (N + 1 < PI) ? N + 1 : PI;
                            static_for<4 * s, KE>([&](auto Kc) {
                                constexpr int K = decltype(Kc)::value;
                                constexpr double f = kFac.inv[N - K];
                                ar = __builtin_fma(f, tri_re(scc, Kc), ar);
                                ai = __builtin_fma(f, tri_im(scc, Kc), ai);
                            });
                            ar = (q + 4 * s > N) ? 0.0 : ar;
                            ai = (q + 4 * s > N) ? 0.0 : ai;
                        }
                        yr[s] = ar;
                        yi[s] = ai;
                    });
                    transform_bwd<N, false, opaque_lane<OP, PI>()>(tab, lane, R, sb, yr, yi);
                    store_sum(ic<N>{}, cr, ci, yr, yi);
                }
                sb *= r;
            }
        });
    } else {
        // L2L: forward scale r^n (translations.cpp:198), backward r^-n
        // (:180), descending degrees.
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
        // Stage n runs the forward transform of source degree n beside the
        // z-translation and backward transform of output degree n + 1
        // (which needs source degrees K >= n + 1 only).
        load_degree(ic<PI - 1>{}, pr, pim);
        static_for_down<-1, PI>([&](auto stc) {
            constexpr int n = decltype(stc)::value;  // forward degree
            constexpr int N = n + 1;                 // output degree
            double xr[NSL], xi[NSL], cr[NSL], ci[NSL];
            if constexpr (n >= 0) {
                take(xr, pr);
                take(xi, pim);
                if constexpr (n > 0) load_degree(ic<n - 1>{}, pr, pim);
            }
            if constexpr (N < PE) {
                take(cr, orr);
                take(ci, oi);
            }
            if constexpr (n >= 0 && n < PE) fetch_old(ic<n>{}, orr, oi);
            sched_fence();
            if constexpr (n >= 0) {
                if (!RT || uint32_t(n) < pin_w) {
                    transform_fwd<n, true, opaque_lane<OP, PI>()>(tab, lane, R, fs[n], xr, xi);
                    keep_degree(ic<n>{}, xr, xi);
                } else {
                    clear_degree(ic<n>{});
                }
            }
            if constexpr (N < PE) {
                if (!RT || uint32_t(N) < pout_w) {
                    double yr[NSL], yi[NSL];
                    static_for<0, NSL>([&](auto scc) {
                        constexpr int s = decltype(scc)::value;
                        double ar = 0.0, ai = 0.0;
                        if constexpr (4 * s <= N && 4 * s < PI) {
                            static_for<N, PI>([&](auto Kc) {
                                constexpr int K = decltype(Kc)::value;
                                constexpr double f = kFac.inv[K - N];
                                ar = __builtin_fma(f, tri_re(scc, Kc), ar);
                                ai = __builtin_fma(f, tri_im(scc, Kc), ai);
                            });
                            ar = (q + 4 * s > N) ? 0.0 : ar;
                            ai = (q + 4 * s > N) ? 0.0 : ai;
                        }
                        yr[s] = ar;
                        yi[s] = ai;
                    });
                    transform_bwd<N, true, opaque_lane<OP, PI>()>(tab, lane, R, bs[N], yr, yi);
                    store_sum(ic<N>{}, cr, ci, yr, yi);
                }
            }
        });
    }

    if (zero_shift) {
        // M2M / L2L with zero shift: target += source over the common
        // degrees (translations.cpp:539-547, 596-604); quarter-wave q takes
        // every fourth real.
        const int len = RT ? int(pend * (pend + 1)) : PE * (PE + 1);
        for (int i = q; i < len; i += 4) D[i] = (acc ? D[i] : 0.0) + S[i];
    }
}

// Persistent workgroups of kWaves waves: stage the swap blocks of degrees
// < max(PI, PO) in LDS, then each wave takes groups of 16 translations.
template <int OP, int PI, int PO, bool RT>
__global__ void __launch_bounds__(64 * kWaves)
    k_translate_mfma(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                     const sfmm_item *__restrict__ items, uint32_t n_items, const double *__restrict__ table, int acc)
{
    constexpr int PM = PI > PO ? PI : PO;
    constexpr int TAB = dense_size(PM);
    __shared__ double tab[TAB];
    constexpr int TL = lds_slot0_size<PI>();
    __shared__ double tri0[TL > 0 ? kWaves * TL : 1];
    for (int i = threadIdx.x; i < TAB; i += 64 * kWaves) tab[i] = table[i];
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    double *tl = tri0 + wave * TL;
    const uint32_t groups = (n_items + 15u) / 16u;
    for (uint32_t g = blockIdx.x * kWaves + wave; g < groups; g += gridDim.x * kWaves) {
        // The table is the same for every group: an opaque zero offset per
        // group keeps the compiler from hoisting all of its reads out of the
        // loop into registers (that would need ~300 registers at order 16);
        // each group reads its operands from LDS where they are used.
        int base = 0;
        asm volatile("" : "+s"(base));
        translate_group<OP, PI, PO, RT>(src_pool, dst_pool, items, n_items, g, tab + base, tl, lane, acc);
    }
}

using fn_t = void (*)(const double *, double *, const sfmm_item *, uint32_t, const double *, int);

template <int OP, bool RT, int... Ps>
constexpr auto make_table(std::integer_sequence<int, Ps...>)
{
    struct T {
        fn_t f[sizeof...(Ps)];
    } t{{&k_translate_mfma<OP, Ps + kMfmaMin, Ps + kMfmaMin, RT>...}};
    return t;
}

// Launch the square-order-P kernel of (OP, RT) over n items as persistent
// workgroups.  One translation unit instantiates each (OP, RT) pair
// (mfma_{m2m,m2l,l2l}{,_rt}.hip) so they compile in parallel.
template <int OP, bool RT>
hipError_t launch_op(uint32_t P, const sfmm_item *items, size_t n, const double *src, double *dst,
                     const double *table, hipStream_t stream, int acc)
{
    static const auto tab = make_table<OP, RT>(std::make_integer_sequence<int, kMfmaMax - kMfmaMin + 1>{});
    if (P < uint32_t(kMfmaMin) || P > uint32_t(kMfmaMax)) return hipErrorInvalidValue;
    const fn_t f = tab.f[P - kMfmaMin];
    const size_t groups = (n + 15) / 16;
    const uint32_t grid =
        persistent_grid(reinterpret_cast<const void *>(f), 64 * kWaves, (groups + kWaves - 1) / kWaves);
    hipLaunchKernelGGL(f, dim3(grid), dim3(64 * kWaves), 0, stream, src, dst, items, static_cast<uint32_t>(n),
                       table, acc);
    return hipGetLastError();
}

}  // namespace mfma
}  // namespace sfmm
