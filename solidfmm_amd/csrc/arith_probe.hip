// Arithmetic-rate probe (sfmm_probe_arith): the rates the reduced-precision
// modes trade against, measured on the device with operands in registers.
//
//   kind 0  FP64 FMA on the VALU (v_fma_f64)
//   kind 1  FP32 FMA on the VALU (v_fma_f32; packed FP32 is not used)
//   kind 2  double-word FP32 multiply-add (prec::DW32::fma: DWTimesDW3 then
//           AccurateDWPlusDW), counted as one multiply-add
//   kind 3  FP64 MFMA 16x16x4 (v_mfma_f64_16x16x4f64, 2048 flop each)
//   kind 4  FP32 MFMA 16x16x4 (v_mfma_f32_16x16x4f32, 2048 flop each)
//
// Every lane runs kChains (8) independent dependency chains, so the rate is
// the issue rate, not the latency; a multiply-add counts two flop.  Eight waves
// per SIMD are launched.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/solidfmm_gpu.h"
#include "prec.hpp"

namespace sfmm {
namespace probe {

constexpr int kChains = 8;
constexpr int kBlock = 256;

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <int KIND>
__global__ void __launch_bounds__(kBlock) k_probe(double *out, int iters, float seed)
{
    const float s = seed + 1e-3f * float(threadIdx.x);
    double r = 0.0;
    if constexpr (KIND == 0) {
        double acc[kChains];
        for (int j = 0; j < kChains; ++j) acc[j] = double(s) + j;
        const double a = 0.999999, b = 1e-7;
        for (int i = 0; i < iters; ++i)
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_fma(acc[j], a, b);
        for (int j = 0; j < kChains; ++j) r += acc[j];
    } else if constexpr (KIND == 1) {
        float acc[kChains];
        for (int j = 0; j < kChains; ++j) acc[j] = s + float(j);
        const float a = 0.99999f, b = 1e-5f;
        for (int i = 0; i < iters; ++i)
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_fmaf(acc[j], a, b);
        for (int j = 0; j < kChains; ++j) r += acc[j];
    } else if constexpr (KIND == 2) {
        using P = prec::DW32;
        prec::dw acc[kChains];
        for (int j = 0; j < kChains; ++j) acc[j] = P::from(double(s) + j);
        const prec::dw a = P::from(0.999999), b = P::from(1e-7);
        for (int i = 0; i < iters; ++i)
            for (int j = 0; j < kChains; ++j) acc[j] = P::fma(acc[j], a, b);
        for (int j = 0; j < kChains; ++j) r += P::to(acc[j]);
    } else if constexpr (KIND == 3) {
        dbl4 acc[kChains];
        for (int j = 0; j < kChains; ++j) acc[j] = dbl4{double(s), 0.0, 0.0, double(j)};
        const double a = 0.5 + 1e-3 * s, b = 0.25;
        for (int i = 0; i < iters; ++i)
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
        for (int j = 0; j < kChains; ++j) r += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    } else {
        flt4 acc[kChains];
        for (int j = 0; j < kChains; ++j) acc[j] = flt4{s, 0.0f, 0.0f, float(j)};
        const float a = 0.5f + 1e-3f * s, b = 0.25f;
        for (int i = 0; i < iters; ++i)
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
        for (int j = 0; j < kChains; ++j) r += double(acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3]);
    }
    out[blockIdx.x * kBlock + threadIdx.x] = r;
}

}  // namespace probe
}  // namespace sfmm

extern "C" int sfmm_probe_arith(int device, int kind, double *flops_per_s)
{
    using namespace sfmm::probe;
    if (!flops_per_s || kind < 0 || kind > 4) return SFMM_ERR_LOGIC;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return SFMM_ERR_HIP;
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const int blocks = (cus > 0 ? cus : 1) * 8;  // 8 waves per SIMD
    const int iters = kind == 2 ? 256 : 4096;
    double *out = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipMalloc(&out, sizeof(double) * size_t(blocks) * kBlock);
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    auto launch = [&](int it) {
        switch (kind) {
        case 0: hipLaunchKernelGGL(k_probe<0>, dim3(blocks), dim3(kBlock), 0, nullptr, out, it, 1.0f); break;
        case 1: hipLaunchKernelGGL(k_probe<1>, dim3(blocks), dim3(kBlock), 0, nullptr, out, it, 1.0f); break;
        case 2: hipLaunchKernelGGL(k_probe<2>, dim3(blocks), dim3(kBlock), 0, nullptr, out, it, 1.0f); break;
        case 3: hipLaunchKernelGGL(k_probe<3>, dim3(blocks), dim3(kBlock), 0, nullptr, out, it, 1.0f); break;
        default: hipLaunchKernelGGL(k_probe<4>, dim3(blocks), dim3(kBlock), 0, nullptr, out, it, 1.0f); break;
        }
        return hipGetLastError();
    };
    if (e == hipSuccess) e = launch(16);  // warm-up
    if (e == hipSuccess) e = hipDeviceSynchronize();
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventRecord(t0, nullptr);
    if (e == hipSuccess) e = launch(iters);
    if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (e == hipSuccess) {
        const double lanes = double(blocks) * kBlock;
        const double per_lane = (kind <= 2) ? 2.0 * kChains : 2048.0 / 64.0 * kChains;
        *flops_per_s = lanes * per_lane * double(iters) / (double(ms) * 1e-3);
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (out) (void)hipFree(out);
    (void)hipSetDevice(prev);
    return e == hipSuccess ? SFMM_OK : SFMM_ERR_HIP;
}
