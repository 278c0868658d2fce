// FP64 translation kernels whose x<->z swap contractions run on the matrix
// cores (v_mfma_f64_16x16x4f64), for orders too large for the one-lane-per-
// translation kernels of translate.hip.
//
// Mapping.  A wave owns 16 translations; lane l = t + 16 q handles
// translation t and the orders m = q + 4 s ("slot" s) of every degree.  That
// is exactly how the MFMA spreads its 16x16 result over the wave (element s
// of lane (q,t) is row q + 4 s, column t), and it is also how the next MFMA
// wants its B operand (rows 4c..4c+3 of K-chunk c sit in element c of the
// four quarter-waves).  So a swap is
//     D[m][t] = sum_c  A_c[m][4c..4c+3] * x[4c..4c+3][t]
// with x taken straight from the previous result registers: no LDS, no
// shuffles between the rotation, swap, rotation, swap steps.  The constant
// A fragments (folded swap blocks, operator_data.cpp:47-147, in MFMA operand
// order) come from a device table built once per process (mfma_tables.cpp).
//
// The rotations e^{+-i m theta}, the z-translation sums and the scaling run
// on the FP64 VALU on the same registers; each lane keeps the rotated
// triangle only for its own orders m = q + 4s, so the z-translation of an
// order m is lane-local.  Algorithm: Dehnen's rotation scheme as in
// translations.cpp:36-471 (see translate.hip for the per-step citations).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "tables.hpp"
#include "translate.hpp"

namespace sfmm {
namespace mfma {

typedef double dbl4 __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One swap step for degree N on the lane's slot values (xr, xi): both
// folded blocks (real and imaginary parts) of variant V (0: the blocks,
// 1: their transposes).
template <int N, int V, int NSL>
__device__ __forceinline__ void swap_mfma(const double *__restrict__ frag, int lane, double (&xr)[NSL],
                                          double (&xi)[NSL])
{
    constexpr int NB = (N + 16) / 16;  // row blocks of 16 orders
    constexpr int NC = (N + 4) / 4;    // K-chunks of 4 orders
    dbl4 dr[NB], di[NB];
    static_for<0, NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        dr[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        di[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        static_for<0, NC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            const double ar = frag[64 * frag_index(V, 0, N, b, c) + lane];
            const double ai = frag[64 * frag_index(V, 1, N, b, c) + lane];
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
}

// Per-lane rotation factors e^{i m alpha}, e^{i m beta} for m = q + 4s.
template <int NSL>
struct LaneRot {
    double ca[NSL], sa[NSL], cb[NSL], sb[NSL];
};

template <int N, int NSL>
__device__ __forceinline__ void rot(const LaneRot<NSL> &R, bool beta, bool forward, double scale, double (&xr)[NSL],
                                    double (&xi)[NSL])
{
    constexpr int NS = (N + 4) / 4;
    static_for<0, NS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const double c = beta ? R.cb[s] : R.ca[s];
        const double sn = beta ? R.sb[s] : R.sa[s];
        const double a = xr[s], b = xi[s];
        if (forward) {
            xr[s] = scale * (c * a - sn * b);
            xi[s] = scale * (sn * a + c * b);
        } else {
            xr[s] = scale * (c * a + sn * b);
            xi[s] = scale * (c * b - sn * a);
        }
    });
}

template <int N, int V, int NSL>
__device__ __forceinline__ void transform_fwd(const double *frag, int lane, const LaneRot<NSL> &R, double scale,
                                              double (&xr)[NSL], double (&xi)[NSL])
{
    rot<N>(R, false, true, scale, xr, xi);
    swap_mfma<N, V>(frag, lane, xr, xi);
    rot<N>(R, true, true, 1.0, xr, xi);
    swap_mfma<N, V>(frag, lane, xr, xi);
}

template <int N, int V, int NSL>
__device__ __forceinline__ void transform_bwd(const double *frag, int lane, const LaneRot<NSL> &R, double scale,
                                              double (&xr)[NSL], double (&xi)[NSL])
{
    swap_mfma<N, V>(frag, lane, xr, xi);
    rot<N>(R, true, false, 1.0, xr, xi);
    swap_mfma<N, V>(frag, lane, xr, xi);
    rot<N>(R, false, false, scale, xr, xi);
}

template <int OP, int PI, int PO>
__global__ void __launch_bounds__(64)
    k_translate_mfma(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                     const sfmm_item *__restrict__ items, uint32_t n_items, const double *__restrict__ frag)
{
    constexpr int PM = PI > PO ? PI : PO;
    constexpr int PE = PI < PO ? PI : PO;
    constexpr int NSL = 4 * ((PM + 15) / 16);  // slots incl. padding to whole row blocks
    const int lane = threadIdx.x;
    const int q = lane >> 4;
    const int t = lane & 15;
    const uint32_t idx = blockIdx.x * 16u + static_cast<uint32_t>(t);
    const bool valid = idx < n_items;

    // Every lane stays active to the end: the MFMAs need the whole wave.
    sfmm_item it{};
    if (valid) it = items[idx];
    const double *__restrict__ S = src_pool + it.src;
    double *__restrict__ D = dst_pool + it.dst;
    double x = it.x, y = it.y, z = it.z;

    bool zero_shift = false;
    if constexpr (OP != SFMM_M2L) zero_shift = valid && (x * x + y * y + z * z == 0.0);
    if constexpr (OP == SFMM_M2M) {
        x = -x;
        y = -y;
        z = -z;
    }
    if (!valid || zero_shift) {  // harmless geometry for idle lanes
        x = 0.0;
        y = 0.0;
        z = 1.0;
    }

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
    LaneRot<NSL> R;
    {
        double c = 1.0, s = 0.0, cb = 1.0, sb = 0.0;
        static_for<0, 4 * NSL>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int slot = k / 4;
            if ((k & 3) == q) {
                R.ca[slot] = c;
                R.sa[slot] = s;
                R.cb[slot] = cb;
                R.sb[slot] = sb;
            }
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

    // Rotated triangle, this lane's orders only.
    double tr[tri_size<PI>()], ti[tri_size<PI>()];
    const bool active = valid && !zero_shift;

    auto load_degree = [&](auto nc, double(&xr)[NSL], double(&xi)[NSL]) {
        constexpr int n = decltype(nc)::value;
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const int m = q + 4 * s;
            xr[s] = 0.0;
            xi[s] = 0.0;
            if constexpr (4 * s <= n) {
                if (active && m <= n) {
                    xr[s] = S[n * (n + 1) + m];
                    xi[s] = S[(n + 1) * (n + 1) + m];
                }
            }
        });
    };
    auto keep_degree = [&](auto nc, const double(&xr)[NSL], const double(&xi)[NSL]) {
        constexpr int n = decltype(nc)::value;
        static_for<0, nslots<PI>()>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (4 * s <= n) {
                tr[slot_off<PI>(s) + n - 4 * s] = xr[s];
                ti[slot_off<PI>(s) + n - 4 * s] = xi[s];
            }
        });
    };
    auto add_degree = [&](auto Nc, const double(&xr)[NSL], const double(&xi)[NSL]) {
        constexpr int N = decltype(Nc)::value;
        static_for<0, NSL>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if constexpr (4 * s <= N) {
                const int m = q + 4 * s;
                if (active && m <= N) {
                    D[N * (N + 1) + m] += xr[s];
                    D[(N + 1) * (N + 1) + m] += xi[s];
                }
            }
        });
    };

    if constexpr (OP == SFMM_M2L) {
        double sc = rinv;  // r^-(n+1) (translations.cpp:45)
        static_for<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            double xr[NSL], xi[NSL];
            load_degree(nc, xr, xi);
            transform_fwd<n, 0>(frag, lane, R, sc, xr, xi);
            keep_degree(nc, xr, xi);
            sc *= rinv;
        });
        double rp[PO];
        rp[0] = 1.0;
        static_for<1, PO>([&](auto Nc) {
            constexpr int N = decltype(Nc)::value;
            rp[N] = rp[N - 1] * rinv;
        });
        static_for_down<0, PO>([&](auto Nc) {
            constexpr int N = decltype(Nc)::value;
            double xr[NSL], xi[NSL];
            const bool odd = ((N + q) & 1) != 0;  // (-1)^(N+m), m = q + 4s
            static_for<0, NSL>([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                double ar = 0.0, ai = 0.0;
                if constexpr (4 * s <= N && 4 * s < PI) {
                    static_for<4 * s, PI>([&](auto Kc) {
                        constexpr int K = decltype(Kc)::value;
                        constexpr double f = kFac.fac[N + K];
                        ar = __builtin_fma(f, tr[slot_off<PI>(s) + K - 4 * s], ar);
                        ai = __builtin_fma(f, ti[slot_off<PI>(s) + K - 4 * s], ai);
                    });
                    if (odd) {
                        ar = -ar;
                        ai = -ai;
                    }
                    if (q + 4 * s > N) {
                        ar = 0.0;
                        ai = 0.0;
                    }
                }
                xr[s] = ar;
                xi[s] = ai;
            });
            transform_bwd<N, 1>(frag, lane, R, rp[N], xr, xi);
            add_degree(Nc, xr, xi);
        });
    } else if constexpr (OP == SFMM_M2M) {
        double sc = rinv;  // forward r^-(n+1)
        double sb = r;     // backward r^(N+1) (translations.cpp:129)
        static_for<0, PM>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            if constexpr (n < PI) {
                double xr[NSL], xi[NSL];
                load_degree(nc, xr, xi);
                transform_fwd<n, 0>(frag, lane, R, sc, xr, xi);
                keep_degree(nc, xr, xi);
                sc *= rinv;
            }
            if constexpr (n < PO) {
                constexpr int N = n;
                double xr[NSL], xi[NSL];
                static_for<0, NSL>([&](auto scc) {
                    constexpr int s = decltype(scc)::value;
                    double ar = 0.0, ai = 0.0;
                    if constexpr (4 * s <= N && 4 * s < PI) {
                        constexpr int KE = (N + 1 < PI) ? N + 1 : PI;
                        static_for<4 * s, KE>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value;
                            constexpr double f = kFac.inv[N - K];
                            ar = __builtin_fma(f, tr[slot_off<PI>(s) + K - 4 * s], ar);
                            ai = __builtin_fma(f, ti[slot_off<PI>(s) + K - 4 * s], ai);
                        });
                        if (q + 4 * s > N) {
                            ar = 0.0;
                            ai = 0.0;
                        }
                    }
                    xr[s] = ar;
                    xi[s] = ai;
                });
                transform_bwd<N, 0>(frag, lane, R, sb, xr, xi);
                add_degree(ic<N>{}, xr, xi);
                sb *= r;
            }
        });
    } else {
        // L2L: forward scale r^n, backward r^-n, descending degrees.
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
        static_for_down<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            {
                double xr[NSL], xi[NSL];
                load_degree(nc, xr, xi);
                transform_fwd<n, 1>(frag, lane, R, fs[n], xr, xi);
                keep_degree(nc, xr, xi);
            }
            if constexpr (n < PE) {
                constexpr int N = n;
                double xr[NSL], xi[NSL];
                static_for<0, NSL>([&](auto scc) {
                    constexpr int s = decltype(scc)::value;
                    double ar = 0.0, ai = 0.0;
                    if constexpr (4 * s <= N && 4 * s < PI) {
                        static_for<N, PI>([&](auto Kc) {
                            constexpr int K = decltype(Kc)::value;
                            constexpr double f = kFac.inv[K - N];
                            ar = __builtin_fma(f, tr[slot_off<PI>(s) + K - 4 * s], ar);
                            ai = __builtin_fma(f, ti[slot_off<PI>(s) + K - 4 * s], ai);
                        });
                        if (q + 4 * s > N) {
                            ar = 0.0;
                            ai = 0.0;
                        }
                    }
                    xr[s] = ar;
                    xi[s] = ai;
                });
                transform_bwd<N, 1>(frag, lane, R, bs[N], xr, xi);
                add_degree(ic<N>{}, xr, xi);
            }
        });
    }

    if (zero_shift) {
        // M2M / L2L with zero shift: target += source over the common
        // degrees (translations.cpp:539-547, 596-604); quarter-wave q takes
        // every fourth real.
        constexpr int len = PE * (PE + 1);
        for (int i = q; i < len; i += 4) D[i] += S[i];
    }
}

}  // namespace mfma
}  // namespace sfmm
