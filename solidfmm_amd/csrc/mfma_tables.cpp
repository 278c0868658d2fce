// Host construction of the dense swap-block table (mfma_tables.hpp) from the
// compile-time folded swap blocks of tables.hpp, and its device copy.
#include "mfma_tables.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "tables.hpp"

namespace sfmm {
namespace mfma {

void build_table(double *out)
{
    static_for<0, kMaxOrder>([&](auto nc) {
        constexpr int n = decltype(nc)::value;
        const int st = dense_stride(n);
        double *sr = out + dense_base(n);
        double *si = sr + dense_part(n);
        for (int m = 0; m <= n; ++m)
            for (int l = 0; l < st; ++l) {
                sr[m * st + l] = (l <= n) ? folded<n, 0>(m, l) : 0.0;
                si[m * st + l] = (l <= n) ? folded<n, 1>(m, l) : 0.0;
            }
    });
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

}  // namespace mfma
}  // namespace sfmm
