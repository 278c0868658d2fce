// Fused rotate / translate-along-z / rotate-back kernels for M2M, M2L and
// L2L on one MI355X (gfx950), FP64.
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
// the item touches memory.  All lanes of a launch share (p_in, p_out), so
// the degree guards are uniform scalar branches.  The swap contractions run
// on the FP64 VALU: on MI355X the FP64 MFMA peak equals the FP64 VALU peak
// (78.6 TFLOP/s each) and the folded swap blocks are half zeros, which a
// VALU contraction skips and an MFMA tile cannot.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "tables.hpp"
#include "translate.hpp"

namespace sfmm {

constexpr int tri(int m, int n) { return n * (n + 1) / 2 + m; }

template <int PB>
struct Rot {
    double ca[PB], sa[PB], cb[PB], sb[PB];
};

// Powers e^{i k alpha}, e^{i k beta} by repeated complex products
// (microkernel_generic.cpp:76-96).
template <int PB>
__device__ __forceinline__ void euler_powers(double ca1, double sa1, double cb1, double sb1, Rot<PB> &R)
{
    R.ca[0] = 1.0;
    R.sa[0] = 0.0;
    R.cb[0] = 1.0;
    R.sb[0] = 0.0;
    if constexpr (PB > 1) {
        R.ca[1] = ca1;
        R.sa[1] = sa1;
        R.cb[1] = cb1;
        R.sb[1] = sb1;
    }
    static_for<2, PB>([&](auto ic_) {
        constexpr int i = decltype(ic_)::value;
        R.ca[i] = R.ca[i - 1] * ca1 - R.sa[i - 1] * sa1;
        R.sa[i] = R.sa[i - 1] * ca1 + R.ca[i - 1] * sa1;
        R.cb[i] = R.cb[i - 1] * cb1 - R.sb[i - 1] * sb1;
        R.sb[i] = R.sb[i - 1] * cb1 + R.cb[i - 1] * sb1;
    });
}

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

// (re,im) <- scale * e^{+i m theta} (re,im)      (rotscale, forward=true)
template <int N>
__device__ __forceinline__ void rot_fwd(const double *c, const double *s, double scale, double (&re)[N + 1],
                                        double (&im)[N + 1])
{
    re[0] *= scale;
    im[0] *= scale;
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double a = re[m], b = im[m];
        re[m] = scale * (c[m] * a - s[m] * b);
        im[m] = scale * (s[m] * a + c[m] * b);
    });
}

template <int N>
__device__ __forceinline__ void rot_fwd1(const double *c, const double *s, double (&re)[N + 1], double (&im)[N + 1])
{
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double a = re[m], b = im[m];
        re[m] = c[m] * a - s[m] * b;
        im[m] = s[m] * a + c[m] * b;
    });
}

// (re,im) <- scale * e^{-i m theta} (re,im)      (rotscale, forward=false)
template <int N>
__device__ __forceinline__ void rot_bwd(const double *c, const double *s, double scale, double (&re)[N + 1],
                                        double (&im)[N + 1])
{
    re[0] *= scale;
    im[0] *= scale;
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double a = re[m], b = im[m];
        re[m] = scale * (c[m] * a + s[m] * b);
        im[m] = scale * (c[m] * b - s[m] * a);
    });
}

template <int N>
__device__ __forceinline__ void rot_bwd1(const double *c, const double *s, double (&re)[N + 1], double (&im)[N + 1])
{
    static_for<1, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double a = re[m], b = im[m];
        re[m] = c[m] * a + s[m] * b;
        im[m] = c[m] * b - s[m] * a;
    });
}

// Issue the loads of degree N of a block into v[0..N] (16-byte pieces).
// Callers place a scheduling barrier after it, so the loads stay where they
// were issued and their latency overlaps the work that follows.
template <int N, int CAP>
__device__ __forceinline__ void fetch_degree(const double *__restrict__ blk, double2 (&v)[CAP])
{
    static_assert(N < CAP, "prefetch buffer too small");
    const double2 *q = reinterpret_cast<const double2 *>(blk + N * (N + 1));
    static_for<0, N + 1>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = q[j];
    });
}

// target degree N = old + (re, im), old prefetched by fetch_degree
// (buf2solid, microkernel_generic.cpp:361-380).
template <int N, int CAP>
__device__ __forceinline__ void add_store_degree(double *__restrict__ blk, const double2 (&old)[CAP],
                                                 const double (&re)[N + 1], const double (&im)[N + 1])
{
    double2 *q = reinterpret_cast<double2 *>(blk + N * (N + 1));
    double a[2 * (N + 1)];
    static_for<0, N + 1>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        a[m] = re[m];
        a[N + 1 + m] = im[m];
    });
    static_for<0, N + 1>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        double2 v = old[j];
        v.x += a[2 * j];
        v.y += a[2 * j + 1];
        q[j] = v;
    });
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

// Forward pipelines: M-type (M2M, M2L; translations.cpp:36-85) uses the
// folded blocks, L-type (L2L; translations.cpp:189-239) their transposes.
template <int N, bool LTYPE, int PB>
__device__ __forceinline__ void forward(const Rot<PB> &R, double scale, double (&re)[N + 1], double (&im)[N + 1])
{
    constexpr int KR = LTYPE ? 2 : 0;
    constexpr int KI = LTYPE ? 3 : 1;
    rot_fwd<N>(R.ca, R.sa, scale, re, im);
    swap_stage<N, KR, KI>(re, im);
    rot_fwd1<N>(R.cb, R.sb, re, im);
    swap_stage<N, KR, KI>(re, im);
}

// Backward pipelines: M-type (M2M; translations.cpp:87-136), L-type
// (M2L, L2L; translations.cpp:138-187).
template <int N, bool LTYPE, int PB>
__device__ __forceinline__ void backward(const Rot<PB> &R, double scale, double (&re)[N + 1], double (&im)[N + 1])
{
    constexpr int KR = LTYPE ? 2 : 0;
    constexpr int KI = LTYPE ? 3 : 1;
    swap_stage<N, KR, KI>(re, im);
    rot_bwd1<N>(R.cb, R.sb, re, im);
    swap_stage<N, KR, KI>(re, im);
    rot_bwd<N>(R.ca, R.sa, scale, re, im);
}

// Zero shift in M2M / L2L: target += source over the common degrees
// (translations.cpp:539-547, 596-604).
__device__ __forceinline__ void zero_shift_add(const double *__restrict__ S, double *__restrict__ D, uint32_t p)
{
    const uint32_t len = p * (p + 1);
    for (uint32_t i = 0; i < len; ++i) D[i] += S[i];
}

// One lane, one translation.  PI/PO are the orders when EXACT; otherwise
// bounds, with the runtime orders pi, po <= PB (uniform over the launch).
template <int OP, int PI, int PO>
__global__ void __launch_bounds__(64)
    k_translate(const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                const sfmm_item *__restrict__ items, uint32_t n_items)
{
    constexpr int PM = PI > PO ? PI : PO;
    constexpr int PE = PI < PO ? PI : PO;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_items) return;

    const sfmm_item it = items[t];
    const double *__restrict__ S = static_cast<const double *>(__builtin_assume_aligned(src_pool + it.src, 16));
    double *__restrict__ D = static_cast<double *>(__builtin_assume_aligned(dst_pool + it.dst, 16));
    double x = it.x, y = it.y, z = it.z;

    if constexpr (OP != SFMM_M2L) {
        if (x * x + y * y + z * z == 0.0) {
            zero_shift_add(S, D, PE);
            return;
        }
    }
    if constexpr (OP == SFMM_M2M) {  // translations.cpp:331-333
        x = -x;
        y = -y;
        z = -z;
    }

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
    const double cb1 = z * rinv, sb1 = -rho * rinv;
    Rot<PM> R;
    euler_powers<PM>(ca1, sa1, cb1, sb1, R);

    // Whole source block in flight at once.  Its registers are handed over
    // degree by degree to the rotated triangle, so this costs no peak.
    double2 sv[tri(0, PI)];
    {
        const double2 *q = reinterpret_cast<const double2 *>(S);
        static_for<0, tri(0, PI)>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            sv[j] = q[j];
        });
    }
    // Target degrees are prefetched one ahead of their use (tg); the
    // scheduling fences keep those loads where they are issued.
    double2 tg[PO];

    // Rotated triangle: T[tri(m,K)] holds order m of degree K.
    double tre[tri(0, PI)], tim[tri(0, PI)];

    if constexpr (OP == SFMM_M2L) {
        // Forward, all input degrees; scale r^-(n+1) (translations.cpp:45).
        double s = rinv;
        static_for<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            if constexpr (n == PI - 1) {
                fetch_degree<PO - 1>(D, tg);
                sched_fence();
            }
            double re[n + 1], im[n + 1];
            unpack_block_degree<n>(sv, re, im);
            forward<n, false, PM>(R, s, re, im);
            static_for<0, n + 1>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                tre[tri(m, n)] = re[m];
                tim[tri(m, n)] = im[m];
            });
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
            static_for<0, N + 1>([&](auto jc) { cur[decltype(jc)::value] = tg[decltype(jc)::value]; });
            if constexpr (N > 0) fetch_degree<N - 1>(D, tg);
            sched_fence();
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
            backward<N, true, PM>(R, rp[N], re, im);
            add_store_degree<N>(D, cur, re, im);
        });
    } else if constexpr (OP == SFMM_M2M) {
        // Ascending degrees; output degree n needs input degrees K <= n
        // only, so each is finished as soon as its inputs exist.
        // M_N^m = sum_{K<=N} r^{N-K}/(N-K)! T_K^m (translations.cpp:363-377).
        double s = rinv;   // forward scale r^-(n+1)
        double sb = r;     // backward scale r^(N+1) (translations.cpp:129)
        fetch_degree<0>(D, tg);
        sched_fence();
        static_for<0, PM>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            double2 cur[PO];
            if constexpr (n < PO) {
                static_for<0, n + 1>([&](auto jc) { cur[decltype(jc)::value] = tg[decltype(jc)::value]; });
                if constexpr (n + 1 < PO) fetch_degree<n + 1>(D, tg);
                sched_fence();
            }
            if constexpr (n < PI) {
                double re[n + 1], im[n + 1];
                unpack_block_degree<n>(sv, re, im);
                forward<n, false, PM>(R, s, re, im);
                static_for<0, n + 1>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    tre[tri(m, n)] = re[m];
                    tim[tri(m, n)] = im[m];
                });
                s *= rinv;
            }
            if constexpr (n < PO) {
                constexpr int N = n;
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
                backward<N, false, PM>(R, sb, re, im);
                add_store_degree<N>(D, cur, re, im);
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
        if constexpr (PE > 0) {
            fetch_degree<PE - 1>(D, tg);
            sched_fence();
        }
        static_for_down<0, PI>([&](auto nc) {
            constexpr int n = decltype(nc)::value;
            {
                double re[n + 1], im[n + 1];
                unpack_block_degree<n>(sv, re, im);
                forward<n, true, PM>(R, fs[n], re, im);
                static_for<0, n + 1>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    tre[tri(m, n)] = re[m];
                    tim[tri(m, n)] = im[m];
                });
            }
            if constexpr (n < PE) {
                constexpr int N = n;
                double2 cur[PO];
                static_for<0, N + 1>([&](auto jc) { cur[decltype(jc)::value] = tg[decltype(jc)::value]; });
                if constexpr (N > 0) fetch_degree<N - 1>(D, tg);
                sched_fence();
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
                backward<N, true, PM>(R, bs[N], re, im);
                add_store_degree<N>(D, cur, re, im);
            }
        });
    }
}

// ---------------------------------------------------------------------------
// Dispatch: (op, p_in, p_out) -> instantiation.
// ---------------------------------------------------------------------------

using kernel_fn = void (*)(const double *, double *, const sfmm_item *, uint32_t);

template <int OP, int PI, int PO>
constexpr kernel_fn kfn()
{
    return &k_translate<OP, PI, PO>;
}

// Square orders 1..kSquare, all three operators.
constexpr int kSquare = 12;

template <int OP, int... Ps>
constexpr void fill_square(kernel_fn *tab, std::integer_sequence<int, Ps...>)
{
    ((tab[Ps] = kfn<OP, Ps + 1, Ps + 1>()), ...);
}

struct SquareTable {
    kernel_fn f[3][kSquare];
};

static SquareTable make_square_table()
{
    SquareTable t{};
    fill_square<SFMM_M2M>(t.f[0], std::make_integer_sequence<int, kSquare>{});
    fill_square<SFMM_M2L>(t.f[1], std::make_integer_sequence<int, kSquare>{});
    fill_square<SFMM_L2L>(t.f[2], std::make_integer_sequence<int, kSquare>{});
    return t;
}

static const SquareTable &square_table()
{
    static const SquareTable t = make_square_table();
    return t;
}

// Orders up to this use the lane kernels under kAuto; the MFMA kernels
// take the rest (the lane kernels spill above it).
constexpr uint32_t kLaneAutoMax = 10;

bool have_kernel(int op, uint32_t p_in, uint32_t p_out, int family)
{
    if (op < 0 || op > 2 || p_in < 1 || p_out < 1) return false;
    const bool square = p_in == p_out;
    switch (family) {
    case kLaneKernels: return square && p_in <= uint32_t(kSquare);
    case kMfmaKernels: return square && p_in >= uint32_t(mfma::kMfmaMin) && p_in <= uint32_t(mfma::kMfmaMax);
    default: return p_in <= uint32_t(mfma::kMaxOrder) && p_out <= uint32_t(mfma::kMaxOrder);
    }
}

size_t max_kernel_order() { return size_t(mfma::kMaxOrder); }

uint32_t persistent_grid(const void *kernel, int block, size_t work)
{
    struct Key {
        const void *k;
        int dev;
    };
    static std::mutex mu;
    static std::vector<std::pair<Key, uint32_t>> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    uint32_t resident = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (auto &e : cache)
            if (e.first.k == kernel && e.first.dev == dev) resident = e.second;
        if (!resident) {
            int per_cu = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
                per_cu = 1;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                cus = 1;
            resident = uint32_t(per_cu) * uint32_t(cus);
            cache.push_back({{kernel, dev}, resident});
        }
    }
    if (work < resident) return uint32_t(work < 1 ? 1 : work);
    return resident;
}

hipError_t launch_translate(int op, uint32_t p_in, uint32_t p_out, const sfmm_item *items, size_t n,
                            const double *src, double *dst, hipStream_t stream, int family)
{
    if (!have_kernel(op, p_in, p_out, family)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const bool square = p_in == p_out;
    const bool lane = family == kLaneKernels ||
                      (family == kAuto && square && p_in <= kLaneAutoMax);
    const bool mfma_k = !lane && (family == kMfmaKernels ||
                                  (family == kAuto && square && p_in <= uint32_t(mfma::kMfmaMax)));
    if (lane) {
        kernel_fn f = square_table().f[op][p_in - 1];
        const uint32_t block = 64;
        const uint32_t grid = static_cast<uint32_t>((n + block - 1) / block);
        hipLaunchKernelGGL(f, dim3(grid), dim3(block), 0, stream, src, dst, items, static_cast<uint32_t>(n));
        return hipGetLastError();
    }
    if (mfma_k) {
        const double *table = mfma::device_table();
        if (!table) return hipErrorOutOfMemory;
        switch (op) {
        case SFMM_M2M: return mfma::launch_m2m(p_in, items, n, src, dst, table, stream);
        case SFMM_M2L: return mfma::launch_m2l(p_in, items, n, src, dst, table, stream);
        default: return mfma::launch_l2l(p_in, items, n, src, dst, table, stream);
        }
    }
    return launch_generic(op, p_in > p_out ? p_in : p_out, items, n, src, dst, stream);
}

}  // namespace sfmm
