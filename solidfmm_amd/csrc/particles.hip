// Particle-side operators on either side of the translations, batched over
// cells, for D-valued expansions (D blocks per cell, D charges per point):
//   P2M  M_{c,d} += sum_{i in c} q_{i,d} R(P_c, p_i - c)
//   L2P  phi_{i,d} += dot(L_{c,d}, R(P_c, p_i - c)),
//        grad_{i,d} += dot(L_{c,d}, dR(...))
// In the reference these are user code over harmonics::R / dR and
// solid::fmadd / dot (harmonics.cpp:31-174, solid.cpp:40-105, the pattern of
// tests/test_m2l.cpp:70-116).
//
// R follows the reference's recurrences: the diagonal R_m^m from
// R_{m-1}^{m-1} (x + iy)/(2m), then down each column R_{m+1}^m = z R_m^m and
// (n+m)(n-m) R_n^m = (2n-1) z R_{n-1}^m - |x|^2 R_{n-2}^m; dR is the same
// recurrence differentiated term by term.
//
// P2M: one wave per cell, one lane per coefficient (n, m): the
// thread walks the cell's particles (a uniform loop) and evaluates its own
// R_n^m along diagonal m and column m, O(n) per particle, so there is no
// reduction across threads and the block is written once, coalesced.
// L2P: one lane per particle; the lane walks the columns with running
// values (O(1) registers) and reads the cell's coefficients at
// wave-uniform addresses.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/solidfmm_gpu.h"
#include "translate.hpp"

namespace sfmm {
namespace {

constexpr int kP2MThreads = 64;
constexpr int kL2PThreads = 64;

// 1/k for the recurrences' divisors (2j on the diagonal, (n+m)(n-m) =
// (n+m) times (n-m) down a column), so the inner loops multiply instead of
// divide.
constexpr int kInvSize = 2 * 88;  // n + m < 2 * 86 for every order the library takes
struct InvTable {
    double v[kInvSize];
};
constexpr InvTable make_inv()
{
    InvTable t{};
    for (int k = 1; k < kInvSize; ++k) t.v[k] = 1.0 / k;
    return t;
}
__constant__ constexpr InvTable kInv = make_inv();

template <int D>
__global__ void __launch_bounds__(kP2MThreads)
    k_p2m(const double *__restrict__ pts, const sfmm_cell *__restrict__ cells, uint32_t n_cells,
          double *__restrict__ pool, int acc)
{
    constexpr int W = 3 + D;  // doubles per point
    for (uint32_t c = blockIdx.x; c < n_cells; c += gridDim.x) {
        const sfmm_cell cell = cells[c];
        const int P = int(cell.order);
        const int ncoef = P * (P + 1) / 2;
        double *__restrict__ M = pool + cell.offset;
        for (int k = threadIdx.x; k < ncoef; k += kP2MThreads) {
            int n = int((sqrt(8.0 * k + 1.0) - 1.0) * 0.5);
            while (n * (n + 1) / 2 > k) --n;
            while ((n + 1) * (n + 2) / 2 <= k) ++n;
            const int m = k - n * (n + 1) / 2;
            double sr[D], si[D];
            for (int d = 0; d < D; ++d) sr[d] = si[d] = 0.0;
            for (uint32_t i = 0; i < cell.count; ++i) {
                const double *p = pts + W * size_t(cell.first + i);
                const double x = p[0] - cell.x, y = p[1] - cell.y, z = p[2] - cell.z;
                // diagonal: R_0^0 = 1, R_1^1 = (x + iy)/2, R_j^j = R_{j-1}^{j-1} (x + iy)/(2j)
                double dr = 1.0, di = 0.0;
                if (m >= 1) {
                    dr = x / 2;
                    di = y / 2;
                }
                for (int j = 2; j <= m; ++j) {
                    const double f = kInv.v[2 * j];
                    const double a = dr, b = di;
                    dr = (a * x - b * y) * f;
                    di = (a * y + b * x) * f;
                }
                double ar = dr, ai = di;  // R_n^m, walking down column m
                if (n > m) {
                    double br = ar, bi = ai;  // R_{n-2}^m, R_{n-1}^m
                    ar = z * ar;
                    ai = z * ai;
                    const double r2 = x * x + y * y + z * z;
                    for (int nn = m + 2; nn <= n; ++nn) {
                        const double f = kInv.v[nn + m] * kInv.v[nn - m];
                        const double tr = (ar * z * (2 * nn - 1) - r2 * br) * f;
                        const double ti = (ai * z * (2 * nn - 1) - r2 * bi) * f;
                        br = ar;
                        bi = ai;
                        ar = tr;
                        ai = ti;
                    }
                }
                for (int d = 0; d < D; ++d) {
                    sr[d] = fma(p[3 + d], ar, sr[d]);
                    si[d] = fma(p[3 + d], ai, si[d]);
                }
            }
            const int kr = n * (n + 1) + m, ki = (n + 1) * (n + 1) + m;
            for (int d = 0; d < D; ++d) {
                double *Md = M + d * P * (P + 1);
                Md[kr] = (acc ? Md[kr] : 0.0) + sr[d];
                Md[ki] = (acc ? Md[ki] : 0.0) + si[d];
            }
        }
    }
}

template <int D>
__global__ void __launch_bounds__(kL2PThreads)
    k_l2p(const double *__restrict__ pts, const sfmm_cell *__restrict__ cells, uint32_t n_cells,
          const double *__restrict__ pool, double *__restrict__ phi, double *__restrict__ grad)
{
    constexpr int W = 3 + D;  // doubles per point
    for (uint32_t c = blockIdx.x; c < n_cells; c += gridDim.x) {
        const sfmm_cell cell = cells[c];
        const int P = int(cell.order);
        const double *__restrict__ L = pool + cell.offset;
        for (uint32_t i0 = 0; i0 < cell.count; i0 += kL2PThreads) {
            const uint32_t i = i0 + threadIdx.x;
            const bool valid = i < cell.count;
            const double *p = pts + W * size_t(cell.first + (valid ? i : 0));
            const double x = p[0] - cell.x, y = p[1] - cell.y, z = p[2] - cell.z;
            const double r2 = x * x + y * y + z * z;
            // running diagonal value and gradient: R_m^m, d/dx, d/dy, d/dz
            double dr = 1.0, di = 0.0, gdr[3] = {0.0, 0.0, 0.0}, gdi[3] = {0.0, 0.0, 0.0};
            double f0[D], g[D][3];
            for (int d = 0; d < D; ++d) f0[d] = g[d][0] = g[d][1] = g[d][2] = 0.0;
            for (int m = 0; m < P; ++m) {
                if (m == 1) {
                    dr = x / 2;
                    di = y / 2;
                    gdr[0] = 0.5;
                    gdi[1] = 0.5;
                } else if (m >= 2) {
                    const double f = kInv.v[2 * m];
                    const double a = dr, b = di;
                    const double ax = gdr[0], ay = gdr[1], az = gdr[2], bx = gdi[0], by = gdi[1], bz = gdi[2];
                    dr = (a * x - b * y) * f;
                    gdr[0] = (ax * x + a - bx * y) * f;
                    gdr[1] = (ay * x - by * y - b) * f;
                    gdr[2] = (az * x - bz * y) * f;
                    di = (a * y + b * x) * f;
                    gdi[0] = (ax * y + bx * x + b) * f;
                    gdi[1] = (ay * y + a + by * x) * f;
                    gdi[2] = (az * y + bz * x) * f;
                }
                // dot weight: 2 for m >= 1, 1 for m == 0 (solid.cpp:51-66)
                const double w = m ? 2.0 : 1.0;
                // column m: a = R_n^m, b = R_{n-1}^m (with gradients)
                double ar = dr, ai = di, gar[3] = {gdr[0], gdr[1], gdr[2]}, gai[3] = {gdi[0], gdi[1], gdi[2]};
                double br = 0.0, bi = 0.0, gbr[3] = {0.0, 0.0, 0.0}, gbi[3] = {0.0, 0.0, 0.0};
                for (int n = m; n < P; ++n) {
                    if (n == m + 1) {  // R_{m+1}^m = z R_m^m
                        br = ar;
                        bi = ai;
                        for (int d = 0; d < 3; ++d) {
                            gbr[d] = gar[d];
                            gbi[d] = gai[d];
                        }
                        ar = z * br;
                        ai = z * bi;
                        for (int d = 0; d < 3; ++d) {
                            gar[d] = z * gbr[d];
                            gai[d] = z * gbi[d];
                        }
                        gar[2] += br;
                        gai[2] += bi;
                    } else if (n >= m + 2) {
                        const double f = kInv.v[n + m] * kInv.v[n - m];
                        const double k2 = 2.0 * n - 1.0;
                        const double tr = (ar * z * k2 - r2 * br) * f;
                        const double ti = (ai * z * k2 - r2 * bi) * f;
                        double tgr[3], tgi[3];
                        for (int d = 0; d < 3; ++d) {
                            const double r2d = d == 0 ? 2 * x : d == 1 ? 2 * y : 2 * z;
                            tgr[d] = gar[d] * z * k2 - r2d * br - r2 * gbr[d];
                            tgi[d] = gai[d] * z * k2 - r2d * bi - r2 * gbi[d];
                        }
                        tgr[2] += ar * k2;
                        tgi[2] += ai * k2;
                        br = ar;
                        bi = ai;
                        ar = tr;
                        ai = ti;
                        for (int d = 0; d < 3; ++d) {
                            gbr[d] = gar[d];
                            gbi[d] = gai[d];
                            gar[d] = tgr[d] * f;
                            gai[d] = tgi[d] * f;
                        }
                    }
                    for (int d = 0; d < D; ++d) {
                        const double *Ld = L + d * P * (P + 1);
                        const double lr = Ld[n * (n + 1) + m], li = Ld[(n + 1) * (n + 1) + m];
                        f0[d] = fma(w * lr, ar, fma(w * li, ai, f0[d]));
                        for (int k = 0; k < 3; ++k) g[d][k] = fma(w * lr, gar[k], fma(w * li, gai[k], g[d][k]));
                    }
                }
            }
            if (valid) {
                const size_t j = cell.first + i;
                for (int d = 0; d < D; ++d) {
                    phi[j * D + d] += f0[d];
                    if (grad) {
                        grad[3 * (j * D + d)] += g[d][0];
                        grad[3 * (j * D + d) + 1] += g[d][1];
                        grad[3 * (j * D + d) + 2] += g[d][2];
                    }
                }
            }
        }
    }
}

}  // namespace

namespace {

template <int D>
hipError_t p2m_d(const double *pts, const sfmm_cell *cells, size_t n_cells, double *pool, hipStream_t stream)
{
    const uint32_t grid = persistent_grid(reinterpret_cast<const void *>(&k_p2m<D>), kP2MThreads, n_cells);
    hipLaunchKernelGGL(k_p2m<D>, dim3(grid), dim3(kP2MThreads), 0, stream, pts, cells, uint32_t(n_cells), pool, 1);
    return hipGetLastError();
}

template <int D>
hipError_t l2p_d(const double *pts, const sfmm_cell *cells, size_t n_cells, const double *pool, double *phi,
                 double *grad, hipStream_t stream)
{
    const uint32_t grid = persistent_grid(reinterpret_cast<const void *>(&k_l2p<D>), kL2PThreads, n_cells);
    hipLaunchKernelGGL(k_l2p<D>, dim3(grid), dim3(kL2PThreads), 0, stream, pts, cells, uint32_t(n_cells), pool, phi,
                       grad);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_p2m(const double *pts, const sfmm_cell *cells, size_t n_cells, double *pool, int dim,
                      hipStream_t stream)
{
    if (n_cells == 0) return hipSuccess;
    if (n_cells > 0xFFFFFFFFull) return hipErrorInvalidValue;
    switch (dim) {
    case 1: return p2m_d<1>(pts, cells, n_cells, pool, stream);
    case 2: return p2m_d<2>(pts, cells, n_cells, pool, stream);
    case 3: return p2m_d<3>(pts, cells, n_cells, pool, stream);
    case 4: return p2m_d<4>(pts, cells, n_cells, pool, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_l2p(const double *pts, const sfmm_cell *cells, size_t n_cells, const double *pool, double *phi,
                      double *grad, int dim, hipStream_t stream)
{
    if (n_cells == 0) return hipSuccess;
    if (n_cells > 0xFFFFFFFFull) return hipErrorInvalidValue;
    switch (dim) {
    case 1: return l2p_d<1>(pts, cells, n_cells, pool, phi, grad, stream);
    case 2: return l2p_d<2>(pts, cells, n_cells, pool, phi, grad, stream);
    case 3: return l2p_d<3>(pts, cells, n_cells, pool, phi, grad, stream);
    case 4: return l2p_d<4>(pts, cells, n_cells, pool, phi, grad, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace sfmm
