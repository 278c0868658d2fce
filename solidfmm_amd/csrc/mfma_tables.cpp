// Host construction of the MFMA operand table (mfma_tables.hpp) from the
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
        static_for<0, 4>([&](auto kc) {
            constexpr int kind = decltype(kc)::value;  // 0,1: blocks re/im; 2,3: transposes
            const int v = kind / 2, p = kind % 2;
            for (int b = 0; b < row_blocks(n); ++b)
                for (int c = 0; c < k_chunks(n); ++c) {
                    double *f = out + size_t(64) * frag_index(v, p, n, b, c);
                    for (int lane = 0; lane < 64; ++lane) {
                        const int m = 16 * b + (lane % 16);
                        const int l = 4 * c + lane / 16;
                        f[lane] = (m <= n && l <= n) ? folded<n, kind>(m, l) : 0.0;
                    }
                }
        });
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
