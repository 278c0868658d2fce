// Reduced-precision translation kernel for any orders (up to the precision
// limit) and any mix of (p_in, p_out) within one launch: a group of 16 lanes
// per translation, four translations per wave (64-thread workgroups), the
// rotated triangle and the working vectors in the group's own LDS region,
// lane j of a group handling the orders m = j, j + 16 of each degree; the
// bulk arithmetic in an arithmetic policy of prec.hpp (FP32, or
// double-word FP32).  A group's lanes exchange values only through LDS
// within one wave, so no workgroup barrier is needed past the set-up.
//
// Scaled basis (as in lane_prec.hip): order m of degree n is carried scaled
// by a power of two 2^q(n,m), with 2^e(n,m) ~ sqrt((n-m)!(n+m)!),
//   M-type sides  q = e(n,m) - e(n,n),  blocks 2^e(n,m) S 2^-e(n,l),
//   L-type sides  q = e(n,0) - e(n,m),  blocks 2^-e(n,m) S^T' 2^e(n,l),
// which makes the swap blocks near-orthogonal (their entries reach ~2e6 at
// degree 23 in solidfmm's normalisation, operator_data.cpp:47-147) and keeps
// every intermediate inside the FP32 exponent range up to order 24.  The
// change of basis is exact.
//
// Geometry and scaling stay FP64: the Euler powers are computed in FP64 and
// converted once; each source degree is scaled in FP64 and then converted;
// each output degree is converted back to FP64, scaled in FP64 and added
// into the FP64 target.  M2L carries K! r^-(K+1) 2^K forward and
// N! r^-N 2^N backward, so its Hankel weight (N+K)! (translations.cpp:
// 283-296) becomes C(N+K, K) 2^-(N+K) <= 1.  M2M and L2L keep the
// reference's scales (translations.cpp:129-132, 180, 198) and Toeplitz
// weights 1/(N-K)!.  Steps otherwise as translations.cpp:36-471.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "prec.hpp"
#include "translate.hpp"

namespace sfmm {
namespace precgen {

constexpr int kThreads = 64;
constexpr int kGroup = 16;                    // lanes per translation
constexpr int kGroups = kThreads / kGroup;    // translations per wave

// Orders LDS writes and reads of one wave's lanes (they run in program
// order on the wave; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int dense_base_rt(int n)
{
    int s = 0;
    for (int j = 0; j < n; ++j) s += 2 * (j + 1) * ((j + 1) | 1);
    return s;
}

// Basis exponent e(n, m) = floor(log2((n-m)! (n+m)!) / 2).
__device__ __forceinline__ int enorm_rt(const double *__restrict__ fac, int n, int m)
{
    return ilogb(fac[n - m] * fac[n + m]) >> 1;
}

// out = S_n in (tr: D^-1 S_n^T D in) for the real and imaginary blocks, in
// the scaled basis: entry (m, l) times 2^(e[m] - e[l]) (M-type) or
// 2^(e[l] - e[m]) (tr, L-type); e holds e(n, .).  Lane j of a group
// computes the rows m = j, j + 16, ...; the real block is zero unless
// l + m + n is even and the imaginary block unless it is odd
// (operator_data.cpp:47-147), so each row visits one parity of l per part
// (skipping exact zeros leaves every sum unchanged).
template <class Pol>
__device__ void swap_p(const double *__restrict__ table, int n, bool tr, const int *e, const typename Pol::T *in_r,
                       const typename Pol::T *in_i, typename Pol::T *out_r, typename Pol::T *out_i, int j)
{
    using T = typename Pol::T;
    const int st = (n + 1) | 1;
    const double *sr = table + dense_base_rt(n);
    const double *si = sr + (n + 1) * st;
    for (int m = j; m <= n; m += kGroup) {
        T ar = Pol::zero(), ai = Pol::zero();
        const int lr = (m + n) & 1;  // first l of the real part's parity
        for (int l = lr; l <= n; l += 2) {
            double cr;
            if (!tr) {
                cr = ldexp(sr[m * st + l], e[m] - e[l]);
            } else {
                const double f = (l == 0 ? 1.0 : 2.0) * (m == 0 ? 1.0 : 0.5);
                cr = ldexp(f * sr[l * st + m], e[l] - e[m]);
            }
            if (cr != 0.0) ar = Pol::fma(Pol::from(cr), in_r[l], ar);
        }
        for (int l = lr ^ 1; l <= n; l += 2) {
            double ci;
            if (!tr) {
                ci = ldexp(si[m * st + l], e[m] - e[l]);
            } else {
                const double f = (l == 0 ? 1.0 : 2.0) * (m == 0 ? 1.0 : 0.5);
                ci = ldexp(f * si[l * st + m], e[l] - e[m]);
            }
            if (ci != 0.0) ai = Pol::fma(Pol::from(ci), in_i[l], ai);
        }
        out_r[m] = ar;
        out_i[m] = ai;
    }
}

// LDS of a workgroup: the exponent triangle (16-byte aligned), then per
// group the triangle (re, im), eight vectors of pc values of T (sized for
// 8-byte T) and two of pc doubles.
__host__ __device__ constexpr size_t eb_bytes(int pc) { return (sizeof(int) * size_t(pc) * (pc + 1) / 2 + 15) / 16 * 16; }
__host__ __device__ constexpr size_t group_bytes(int pc)
{
    return 8 * (size_t(pc) * (pc + 1) + 8 * size_t(pc)) + 2 * sizeof(double) * size_t(pc);
}

template <class Pol>
__global__ void __launch_bounds__(kThreads)
    k_translate_prec(int op, const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                     const sfmm_item *__restrict__ items, uint32_t n_items, const double *__restrict__ table,
                     const double *__restrict__ fac, const double *__restrict__ invfac, int pc, int acc)
{
    using T = typename Pol::T;
    extern __shared__ double sh_raw[];
    const int tri = pc * (pc + 1) / 2;
    const int tid = threadIdx.x;
    const int j = tid % kGroup;  // lane within the group
    // e(n, m) for the triangle, shared by the groups; then one region per
    // group (group_bytes)
    int *eb = reinterpret_cast<int *>(sh_raw);
    unsigned char *gbase = reinterpret_cast<unsigned char *>(sh_raw) + eb_bytes(pc) + (tid / kGroup) * group_bytes(pc);
    T *tre = reinterpret_cast<T *>(gbase);
    T *tim = tre + tri;
    T *var = tim + tri;
    T *vai = var + pc;
    T *vbr = vai + pc;
    T *vbi = vbr + pc;
    T *ca = vbi + pc;  // e^{i k alpha}, e^{i k beta}, k < pc, in T
    T *sa = ca + pc;
    T *cb = sa + pc;
    T *sb = cb + pc;
    double *fs = reinterpret_cast<double *>(gbase + 8 * (2 * tri + 8 * pc));  // forward scale per degree
    double *bs = fs + pc;                                                    // backward scale per degree
    for (int n = 0; n < pc; ++n)
        for (int m = tid; m <= n; m += kThreads) eb[n * (n + 1) / 2 + m] = enorm_rt(fac, n, m);
    __syncthreads();
    // basis exponents: forward (source side) and backward (target side)
    auto qside = [&](bool ltype, int n, int m) {
        const int *e = eb + n * (n + 1) / 2;
        return ltype ? e[0] - e[m] : e[m] - e[n];
    };

    for (uint32_t item = blockIdx.x * kGroups + tid / kGroup; item < n_items; item += gridDim.x * kGroups) {
        const sfmm_item it = items[item];
        const int pi = int(it.p_in), po = int(it.p_out);
        if (pi == 0 || po == 0) continue;
        const double *__restrict__ S = src_pool + it.src;
        double *__restrict__ D = dst_pool + it.dst;
        double x = it.x, y = it.y, z = it.z;
        const int pend = pi < po ? pi : po;
        if (op != SFMM_M2L && x * x + y * y + z * z == 0.0) {  // translations.cpp:539-547, 596-604
            for (int i = j; i < pend * (pend + 1); i += kGroup) D[i] = (acc ? D[i] : 0.0) + S[i];
            continue;
        }
        if (op == SFMM_M2M) {
            x = -x;
            y = -y;
            z = -z;
        }
        const int pm = pi > po ? pi : po;
        if (j == 0) {
            // Euler data and powers in FP64 (microkernel_generic.cpp:36-97).
            const double rho = sqrt(x * x + y * y);
            const double r = sqrt(x * x + y * y + z * z);
            const double rinv = 1.0 / r;
            const double c1 = (rho == 0.0) ? 1.0 : y / rho, s1 = (rho == 0.0) ? 0.0 : x / rho;
            const double cb1 = z * rinv, sb1 = -rho * rinv;
            double pca = 1.0, psa = 0.0, pcb = 1.0, psb = 0.0;
            double rp = 1.0, rm = 1.0;  // r^k, r^-k
            for (int k = 0; k < pm; ++k) {
                ca[k] = Pol::from(pca);
                sa[k] = Pol::from(psa);
                cb[k] = Pol::from(pcb);
                sb[k] = Pol::from(psb);
                const double na = pca * c1 - psa * s1, nsa = psa * c1 + pca * s1;
                const double nb = pcb * cb1 - psb * sb1, nsb = psb * cb1 + pcb * sb1;
                pca = na;
                psa = nsa;
                pcb = nb;
                psb = nsb;
                if (op == SFMM_M2L) {
                    fs[k] = ldexp(fac[k] * rm * rinv, k);  // k! r^-(k+1) 2^k
                    bs[k] = ldexp(fac[k] * rm, k);         // k! r^-k 2^k
                } else if (op == SFMM_M2M) {
                    fs[k] = rm * rinv;  // r^-(k+1)
                    bs[k] = rp * r;     // r^(k+1)
                } else {
                    fs[k] = rp;  // r^k
                    bs[k] = rm;  // r^-k
                }
                rp *= r;
                rm *= rinv;
            }
        }
        wave_sync();

        // Forward transforms: M-type for M2M / M2L, L-type for L2L.
        const bool ltype_fwd = (op == SFMM_L2L);
        for (int n = 0; n < pi; ++n) {
            const double scale = fs[n];
            const int *en = eb + n * (n + 1) / 2;
            for (int m = j; m <= n; m += kGroup) {
                const int q = qside(ltype_fwd, n, m);
                const T a = Pol::from(ldexp(scale * S[n * (n + 1) + m], q));
                const T b = Pol::from(ldexp(scale * S[(n + 1) * (n + 1) + m], q));
                var[m] = Pol::sub(Pol::mul(ca[m], a), Pol::mul(sa[m], b));
                vai[m] = Pol::add(Pol::mul(sa[m], a), Pol::mul(ca[m], b));
            }
            wave_sync();
            swap_p<Pol>(table, n, ltype_fwd, en, var, vai, vbr, vbi, j);
            wave_sync();
            for (int m = j; m <= n; m += kGroup) {
                const T a = vbr[m], b = vbi[m];
                var[m] = Pol::sub(Pol::mul(cb[m], a), Pol::mul(sb[m], b));
                vai[m] = Pol::add(Pol::mul(sb[m], a), Pol::mul(cb[m], b));
            }
            wave_sync();
            swap_p<Pol>(table, n, ltype_fwd, en, var, vai, vbr, vbi, j);
            for (int m = j; m <= n; m += kGroup) {
                tre[n * (n + 1) / 2 + m] = vbr[m];
                tim[n * (n + 1) / 2 + m] = vbi[m];
            }
            wave_sync();
        }

        // z-translation and backward transforms: L-type for M2L / L2L,
        // M-type for M2M.
        const bool ltype_bwd = (op != SFMM_M2M);
        const int nend = (op == SFMM_L2L) ? pend : po;
        const int mlim = (op == SFMM_L2L) ? pend : pi;
        for (int N = 0; N < nend; ++N) {
            const int *eN = eb + N * (N + 1) / 2;
            for (int m = j; m <= N; m += kGroup) {
                T ar = Pol::zero(), ai = Pol::zero();
                const int qN = qside(ltype_bwd, N, m);
                if (m < mlim) {
                    if (op == SFMM_M2L) {
                        for (int K = m; K < pi; ++K) {
                            // C(N+K, K) 2^-(N+K), times the basis factors
                            const double c = fac[N + K] * invfac[N] * invfac[K];
                            const T w = Pol::from(ldexp(c, qN - qside(ltype_fwd, K, m) - N - K));
                            ar = Pol::fma(w, tre[K * (K + 1) / 2 + m], ar);
                            ai = Pol::fma(w, tim[K * (K + 1) / 2 + m], ai);
                        }
                        if ((N + m) & 1) {
                            ar = Pol::neg(ar);
                            ai = Pol::neg(ai);
                        }
                    } else if (op == SFMM_M2M) {
                        const int ke = (N + 1 < pi) ? N + 1 : pi;
                        for (int K = m; K < ke; ++K) {
                            const T w = Pol::from(ldexp(invfac[N - K], qN - qside(ltype_fwd, K, m)));
                            ar = Pol::fma(w, tre[K * (K + 1) / 2 + m], ar);
                            ai = Pol::fma(w, tim[K * (K + 1) / 2 + m], ai);
                        }
                    } else {
                        for (int K = N; K < pi; ++K) {
                            const T w = Pol::from(ldexp(invfac[K - N], qN - qside(ltype_fwd, K, m)));
                            ar = Pol::fma(w, tre[K * (K + 1) / 2 + m], ar);
                            ai = Pol::fma(w, tim[K * (K + 1) / 2 + m], ai);
                        }
                    }
                }
                var[m] = ar;
                vai[m] = ai;
            }
            wave_sync();
            swap_p<Pol>(table, N, ltype_bwd, eN, var, vai, vbr, vbi, j);
            wave_sync();
            for (int m = j; m <= N; m += kGroup) {
                const T a = vbr[m], b = vbi[m];
                var[m] = Pol::add(Pol::mul(cb[m], a), Pol::mul(sb[m], b));
                vai[m] = Pol::sub(Pol::mul(cb[m], b), Pol::mul(sb[m], a));
            }
            wave_sync();
            swap_p<Pol>(table, N, ltype_bwd, eN, var, vai, vbr, vbi, j);
            const double scale = bs[N];
            for (int m = j; m <= N; m += kGroup) {
                const T a = vbr[m], b = vbi[m];
                const T re = Pol::add(Pol::mul(ca[m], a), Pol::mul(sa[m], b));
                const T im = Pol::sub(Pol::mul(ca[m], b), Pol::mul(sa[m], a));
                const int q = -qside(ltype_bwd, N, m);
                const double vr = scale * ldexp(Pol::to(re), q);
                const double vi = scale * ldexp(Pol::to(im), q);
                D[N * (N + 1) + m] = (acc ? D[N * (N + 1) + m] : 0.0) + vr;
                D[(N + 1) * (N + 1) + m] = (acc ? D[(N + 1) * (N + 1) + m] : 0.0) + vi;
            }
            wave_sync();
        }
    }
}

size_t lds_bytes(int pc) { return eb_bytes(pc) + kGroups * group_bytes(pc); }

template <class Pol>
hipError_t launch(int op, uint32_t pc, const sfmm_item *items, size_t n, const double *src, double *dst,
                  hipStream_t stream, int acc)
{
    const double *table = mfma::device_table();
    const double *fac = mfma::device_factorials();
    if (!table || !fac) return hipErrorOutOfMemory;
    const size_t lds = lds_bytes(int(pc));
    const uint32_t grid = persistent_grid(reinterpret_cast<const void *>(&k_translate_prec<Pol>), kThreads,
                                          (n + kGroups - 1) / kGroups);
    hipLaunchKernelGGL(k_translate_prec<Pol>, dim3(grid), dim3(kThreads), lds, stream, op, src, dst, items,
                       static_cast<uint32_t>(n), table, fac, fac + 2 * mfma::kMaxOrder, int(pc), acc);
    return hipGetLastError();
}

}  // namespace precgen

hipError_t launch_prec_generic(int precision, int op, uint32_t pc, const sfmm_item *items, size_t n,
                               const double *src, double *dst, hipStream_t stream, int acc)
{
    if (op < 0 || op > 2 || pc < 1 || pc > kPrecMaxOrder) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    switch (precision) {
    case SFMM_FP32: return precgen::launch<prec::F32>(op, pc, items, n, src, dst, stream, acc);
    case SFMM_SPLIT32: return precgen::launch<prec::DW32>(op, pc, items, n, src, dst, stream, acc);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace sfmm
