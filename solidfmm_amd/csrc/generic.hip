// Generic FP64 translation kernel: any orders up to 86 and any mix of
// (p_in, p_out) within one launch.  One workgroup of 64 threads per
// translation (grid-stride over the list); the threads share the order
// index m, the rotated triangle lives in LDS, the folded swap blocks come
// from the dense device table (mfma_tables.hpp).  This is the complete-
// coverage path behind the specialised kernels of translate.hip and
// translate_mfma.hpp, which it must agree with.
//
// Steps follow translations.cpp:36-471: forward transforms per input degree
// (scale, rotate alpha, swap, rotate beta, swap), the z-translation
// (Hankel (N+K)! for M2L, Toeplitz 1/(N-K)! for M2M and L2L), backward
// transforms per output degree and accumulation into the target.  Transposed
// swaps use TR = D^-1 SR^T D, D = diag(1,2,...,2).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"
#include "translate.hpp"

namespace sfmm {
namespace generic {

constexpr int kThreads = 64;

__device__ __forceinline__ int dense_base_rt(int n)
{
    int s = 0;
    for (int j = 0; j < n; ++j) s += 2 * (j + 1) * ((j + 1) | 1);
    return s;
}

// out = S_n in (TR: D^-1 S_n^T D in) for the real and imaginary blocks.
__device__ void swap_g(const double *__restrict__ table, int n, bool tr, const double *in_r, const double *in_i,
                       double *out_r, double *out_i)
{
    const int st = (n + 1) | 1;
    const double *sr = table + dense_base_rt(n);
    const double *si = sr + (n + 1) * st;
    for (int m = threadIdx.x; m <= n; m += kThreads) {
        double ar = 0.0, ai = 0.0;
        for (int l = 0; l <= n; ++l) {
            if (!tr) {
                ar = fma(sr[m * st + l], in_r[l], ar);
                ai = fma(si[m * st + l], in_i[l], ai);
            } else {
                const double f = (l == 0 ? 1.0 : 2.0) * (m == 0 ? 1.0 : 0.5);
                ar = fma(f * sr[l * st + m], in_r[l], ar);
                ai = fma(f * si[l * st + m], in_i[l], ai);
            }
        }
        out_r[m] = ar;
        out_i[m] = ai;
    }
}

__global__ void __launch_bounds__(kThreads)
    k_translate_generic(int op, const double *__restrict__ src_pool, double *__restrict__ dst_pool,
                        const sfmm_item *__restrict__ items, uint32_t n_items, const double *__restrict__ table,
                        const double *__restrict__ fac, const double *__restrict__ invfac, int pc, int acc)
{
    extern __shared__ double sh[];
    const int tri = pc * (pc + 1) / 2;
    double *tre = sh;
    double *tim = tre + tri;
    double *var = tim + tri;
    double *vai = var + pc;
    double *vbr = vai + pc;
    double *vbi = vbr + pc;
    double *ca = vbi + pc;  // e^{i k alpha}, e^{i k beta}, k <= pc
    double *sa = ca + (pc + 1);
    double *cb = sa + (pc + 1);
    double *sb = cb + (pc + 1);
    double *rp = sb + (pc + 1);  // r^k, k <= pc + 1
    double *rm = rp + (pc + 2);  // r^-k
    const int tid = threadIdx.x;

    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const sfmm_item it = items[item];
        const int pi = int(it.p_in), po = int(it.p_out);
        if (pi == 0 || po == 0) continue;
        const double *__restrict__ S = src_pool + it.src;
        double *__restrict__ D = dst_pool + it.dst;
        double x = it.x, y = it.y, z = it.z;
        const int pend = pi < po ? pi : po;
        if (op != SFMM_M2L && x * x + y * y + z * z == 0.0) {  // translations.cpp:539-547, 596-604
            for (int i = tid; i < pend * (pend + 1); i += kThreads) D[i] = (acc ? D[i] : 0.0) + S[i];
            continue;
        }
        if (op == SFMM_M2M) {
            x = -x;
            y = -y;
            z = -z;
        }
        // Euler data and powers (microkernel_generic.cpp:36-97), tables as
        // far as they are read.
        const int kmax = (pi > po ? pi : po) + 2;
        if (tid == 0) {
            const double rho = sqrt(x * x + y * y);
            const double r = sqrt(x * x + y * y + z * z);
            const double rinv = 1.0 / r;
            const double c1 = (rho == 0.0) ? 1.0 : y / rho, s1 = (rho == 0.0) ? 0.0 : x / rho;
            const double cb1 = z * rinv, sb1 = -rho * rinv;
            ca[0] = 1.0; sa[0] = 0.0; cb[0] = 1.0; sb[0] = 0.0;
            rp[0] = 1.0; rm[0] = 1.0;
            rp[1] = r; rm[1] = rinv;
            for (int k = 1; k < kmax && k <= pc; ++k) {
                if (k == 1) {
                    ca[1] = c1; sa[1] = s1; cb[1] = cb1; sb[1] = sb1;
                } else {
                    ca[k] = ca[k - 1] * c1 - sa[k - 1] * s1;
                    sa[k] = sa[k - 1] * c1 + ca[k - 1] * s1;
                    cb[k] = cb[k - 1] * cb1 - sb[k - 1] * sb1;
                    sb[k] = sb[k - 1] * cb1 + cb[k - 1] * sb1;
                }
            }
            for (int k = 2; k <= kmax && k <= pc + 1; ++k) {
                rp[k] = rp[1] * rp[k - 1];
                rm[k] = rm[1] * rm[k - 1];
            }
        }
        __syncthreads();

        // Forward transforms: M-type for M2M / M2L, L-type for L2L.
        const bool ltype_fwd = (op == SFMM_L2L);
        for (int n = 0; n < pi; ++n) {
            const double scale = ltype_fwd ? rp[n] : rm[n + 1];
            for (int m = tid; m <= n; m += kThreads) {
                const double a = S[n * (n + 1) + m], b = S[(n + 1) * (n + 1) + m];
                var[m] = scale * (ca[m] * a - sa[m] * b);
                vai[m] = scale * (sa[m] * a + ca[m] * b);
            }
            __syncthreads();
            swap_g(table, n, ltype_fwd, var, vai, vbr, vbi);
            __syncthreads();
            for (int m = tid; m <= n; m += kThreads) {
                const double a = vbr[m], b = vbi[m];
                var[m] = cb[m] * a - sb[m] * b;
                vai[m] = sb[m] * a + cb[m] * b;
            }
            __syncthreads();
            swap_g(table, n, ltype_fwd, var, vai, vbr, vbi);
            for (int m = tid; m <= n; m += kThreads) {
                tre[n * (n + 1) / 2 + m] = vbr[m];
                tim[n * (n + 1) / 2 + m] = vbi[m];
            }
            __syncthreads();
        }

        // z-translation and backward transforms: L-type for M2L / L2L,
        // M-type for M2M.
        const bool ltype_bwd = (op != SFMM_M2M);
        const int nend = (op == SFMM_L2L) ? pend : po;
        const int mlim = (op == SFMM_L2L) ? pend : pi;
        for (int N = 0; N < nend; ++N) {
            for (int m = tid; m <= N; m += kThreads) {
                double ar = 0.0, ai = 0.0;
                if (m < mlim) {
                    if (op == SFMM_M2L) {
                        for (int K = m; K < pi; ++K) {
                            ar = fma(fac[N + K], tre[K * (K + 1) / 2 + m], ar);
                            ai = fma(fac[N + K], tim[K * (K + 1) / 2 + m], ai);
                        }
                        if ((N + m) & 1) {
                            ar = -ar;
                            ai = -ai;
                        }
                    } else if (op == SFMM_M2M) {
                        const int ke = (N + 1 < pi) ? N + 1 : pi;
                        for (int K = m; K < ke; ++K) {
                            ar = fma(invfac[N - K], tre[K * (K + 1) / 2 + m], ar);
                            ai = fma(invfac[N - K], tim[K * (K + 1) / 2 + m], ai);
                        }
                    } else {
                        for (int K = N; K < pi; ++K) {
                            ar = fma(invfac[K - N], tre[K * (K + 1) / 2 + m], ar);
                            ai = fma(invfac[K - N], tim[K * (K + 1) / 2 + m], ai);
                        }
                    }
                }
                var[m] = ar;
                vai[m] = ai;
            }
            __syncthreads();
            swap_g(table, N, ltype_bwd, var, vai, vbr, vbi);
            __syncthreads();
            for (int m = tid; m <= N; m += kThreads) {
                const double a = vbr[m], b = vbi[m];
                var[m] = cb[m] * a + sb[m] * b;
                vai[m] = cb[m] * b - sb[m] * a;
            }
            __syncthreads();
            swap_g(table, N, ltype_bwd, var, vai, vbr, vbi);
            const double scale = (op == SFMM_M2M) ? rp[N + 1] : rm[N];
            for (int m = tid; m <= N; m += kThreads) {
                const double a = vbr[m], b = vbi[m];
                const double re = scale * (ca[m] * a + sa[m] * b);
                const double im = scale * (ca[m] * b - sa[m] * a);
                D[N * (N + 1) + m] = (acc ? D[N * (N + 1) + m] : 0.0) + re;
                D[(N + 1) * (N + 1) + m] = (acc ? D[(N + 1) * (N + 1) + m] : 0.0) + im;
            }
            __syncthreads();
        }
    }
}

size_t lds_bytes(int pc)
{
    const size_t tri = size_t(pc) * (pc + 1) / 2;
    return sizeof(double) * (2 * tri + 4 * size_t(pc) + 4 * size_t(pc + 1) + 2 * size_t(pc + 2));
}

}  // namespace generic

hipError_t launch_generic(int op, uint32_t pc, const sfmm_item *items, size_t n, const double *src, double *dst,
                          hipStream_t stream, int acc)
{
    if (op < 0 || op > 2 || pc < 1 || pc > uint32_t(mfma::kMaxOrder)) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const double *table = mfma::device_table();
    const double *fac = mfma::device_factorials();
    if (!table || !fac) return hipErrorOutOfMemory;
    const size_t lds = generic::lds_bytes(int(pc));
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&generic::k_translate_generic),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(generic::lds_bytes(mfma::kMaxOrder)));
        attr_set = true;
    }
    const uint32_t grid = persistent_grid(reinterpret_cast<const void *>(&generic::k_translate_generic),
                                          generic::kThreads, n);
    hipLaunchKernelGGL(generic::k_translate_generic, dim3(grid), dim3(generic::kThreads), lds, stream, op, src, dst,
                       items, static_cast<uint32_t>(n), table, fac, fac + 2 * mfma::kMaxOrder, int(pc), acc);
    return hipGetLastError();
}

}  // namespace sfmm
