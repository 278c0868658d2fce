// Host construction of the folded swap blocks for any degree, and of the
// dense device table (layout in mfma_tables.hpp) that the MFMA and generic
// kernels read.
//
// B_n^{m,l} by Dehnen's recurrence (swap_matrix.cpp:148-171): B_0 = 1; the
// row m = 0 of degree n from 0.5 (b(0,l-1) - b(0,l+1)) of degree n-1; rows
// m+1 and -m-1 from rows m and -m of degree n-1.  The folded real and
// imaginary blocks over the stored orders m, l >= 0 follow
// operator_data.cpp:47-147.  Same double operations in the same order as
// the compile-time tables of tables.hpp, so the values agree bit for bit
// (checked by the test-suite through sfmm_swap_block).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/solidfmm_gpu.h"
#include "mfma_tables.hpp"

namespace sfmm {

namespace {

// B_n for n < nmax, each as a (2n+1)^2 array indexed [(m+n) + (l+n)(2n+1)].
struct SwapRecurrence {
    std::vector<std::vector<double>> B;

    double get(int n, int m, int l) const
    {
        if (m < -n || m > n || l < -n || l > n) return 0.0;
        return B[n][(m + n) + (l + n) * (2 * n + 1)];
    }

    void extend(int nmax)
    {
        for (int n = int(B.size()); n < nmax; ++n) {
            std::vector<double> a(size_t(2 * n + 1) * (2 * n + 1), 0.0);
            auto at = [n](int m, int l) { return size_t((m + n) + (l + n) * (2 * n + 1)); };
            if (n == 0) {
                a[0] = 1.0;
            } else {
                for (int l = -n; l <= n; ++l) a[at(0, l)] = 0.5 * (get(n - 1, 0, l - 1) - get(n - 1, 0, l + 1));
                for (int m = 0; m < n; ++m)
                    for (int l = -n; l <= n; ++l) {
                        a[at(m + 1, l)] = 0.5 * (get(n - 1, m, l - 1) + get(n - 1, m, l + 1)) + get(n - 1, m, l);
                        a[at(-m - 1, l)] =
                            0.5 * (get(n - 1, -m, l - 1) + get(n - 1, -m, l + 1)) - get(n - 1, -m, l);
                    }
            }
            B.push_back(std::move(a));
        }
    }

    // Folded block entry [m][l] of degree n; part 0 real, 1 imaginary.
    double folded(int n, int part, int m, int l) const
    {
        const bool odd = ((l + m + n) & 1) != 0;
        if (part == 0 ? odd : !odd) return 0.0;
        const double bp = get(n, l, m), bm = get(n, -l, m);
        if (part == 0) {
            if (l == 0) return get(n, 0, m);
            return (l & 1) ? bp - bm : bp + bm;
        }
        return (l & 1) ? bp + bm : bp - bm;
    }
};

// Built once for every tabulated degree, then only read.
const SwapRecurrence &recurrence()
{
    static const SwapRecurrence rec = [] {
        SwapRecurrence r;
        r.extend(mfma::kMaxOrder);
        return r;
    }();
    return rec;
}

}  // namespace

namespace mfma {

void build_table(double *out)
{
    const SwapRecurrence &rec = recurrence();
    for (int n = 0; n < kMaxOrder; ++n) {
        const int st = dense_stride(n);
        double *sr = out + dense_base(n);
        double *si = sr + dense_part(n);
        for (int m = 0; m <= n; ++m)
            for (int l = 0; l < st; ++l) {
                sr[size_t(m) * st + l] = (l <= n) ? rec.folded(n, 0, m, l) : 0.0;
                si[size_t(m) * st + l] = (l <= n) ? rec.folded(n, 1, m, l) : 0.0;
            }
    }
}

namespace {
struct DeviceTables {
    std::mutex mu;
    std::vector<double *> per_device;
};
DeviceTables &tables_state()
{
    static DeviceTables s;
    return s;
}
}  // namespace

const double *device_table()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    DeviceTables &st = tables_state();
    std::lock_guard<std::mutex> lock(st.mu);
    if (st.per_device.size() <= size_t(dev)) st.per_device.resize(dev + 1, nullptr);
    if (!st.per_device[dev]) {
        std::vector<double> host(table_doubles());
        build_table(host.data());
        double *d = nullptr;
        if (hipMalloc(&d, host.size() * sizeof(double)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            return nullptr;
        }
        st.per_device[dev] = d;
    }
    return st.per_device[dev];
}

const double *device_factorials()
{
    static std::mutex mu;
    static std::vector<double *> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (per_device.size() <= size_t(dev)) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev]) {
        std::vector<double> h(4 * size_t(kMaxOrder));
        h[0] = 1.0;
        // (2 kMaxOrder - 1)! overflows to inf; the kernels read n! only for
        // n <= 2 kMaxOrder - 2.
        for (int i = 1; i < 2 * kMaxOrder; ++i) h[i] = double(i) * h[i - 1];
        for (int i = 0; i < 2 * kMaxOrder; ++i) h[2 * kMaxOrder + i] = 1.0 / h[i];
        double *d = nullptr;
        if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            return nullptr;
        }
        per_device[dev] = d;
    }
    return per_device[dev];
}

}  // namespace mfma
}  // namespace sfmm

extern "C" int sfmm_swap_block(int n, int part, double *out)
{
    if (n < 0 || n >= sfmm::mfma::kMaxOrder || (part != 0 && part != 1) || !out) return SFMM_ERR_LOGIC;
    const auto &rec = sfmm::recurrence();
    for (int m = 0; m <= n; ++m)
        for (int l = 0; l <= n; ++l) out[size_t(m) * (n + 1) + l] = rec.folded(n, part, m, l);
    return SFMM_OK;
}
