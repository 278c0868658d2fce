// Arithmetic policies of the reduced-precision translation kernels.
//
// The bulk arithmetic (rotations, swaps, z-translation) runs in the policy's
// type T; the geometry (Euler angles, their powers, the r^n scale tables)
// is computed in FP64 and converted once, and the final accumulation into
// the FP64 target is FP64.
//
//   F32   T = float.  Round-to-nearest FP32, FMA contractions.
//   DW32  T = dw, a double-word (hi, lo) pair of floats with |lo| <=
//         ulp(hi)/2: about 48 significant bits from FP32 operations only.
//         Sum and product are the algorithms of Joldes, Muller and Popescu,
//         "Tight and rigorous error bounds for basic building blocks of
//         double-word arithmetic", ACM TOMS 44(2), 2017: AccurateDWPlusDW
//         (Algorithm 6, relative error <= 3u^2 + 13u^3) and DWTimesDW3
//         (Algorithm 12, relative error <= 4u^2), built on Knuth's 2Sum,
//         Dekker's Fast2Sum and the FMA-based 2Prod; u = 2^-24.
//
// Both need IEEE round-to-nearest FP32 without reassociation or contraction:
// the library is compiled without fast-math, and every double-word routine
// switches FP contraction off (a contracted Fast2Sum, e.g. fma(xh, yh, cl)
// in place of ch + cl, counts the low part of the product twice).
#pragma once

#include <hip/hip_runtime.h>

namespace sfmm {
namespace prec {

struct dw {
    float hi, lo;
};

struct F32 {
    using T = float;
    static constexpr const char *name = "fp32";
    __device__ static __forceinline__ T from(double v) { return static_cast<float>(v); }
    __device__ static __forceinline__ double to(T v) { return static_cast<double>(v); }
    __device__ static __forceinline__ T zero() { return 0.0f; }
    __device__ static __forceinline__ T add(T a, T b) { return a + b; }
    __device__ static __forceinline__ T sub(T a, T b) { return a - b; }
    __device__ static __forceinline__ T neg(T a) { return -a; }
    __device__ static __forceinline__ T mul(T a, T b) { return a * b; }
    __device__ static __forceinline__ T fma(T a, T b, T c) { return __builtin_fmaf(a, b, c); }
    // a * b with b a power of two (exact scaling)
    __device__ static __forceinline__ T scale2(T a, float b) { return a * b; }
};

struct DW32 {
    using T = dw;
    static constexpr const char *name = "split-fp32";

    // s + e = a + b exactly (Knuth).
    __device__ static __forceinline__ void two_sum(float a, float b, float &s, float &e)
    {
#pragma clang fp contract(off)
        s = a + b;
        const float ap = s - b;
        const float bp = s - ap;
        e = (a - ap) + (b - bp);
    }
    // s + e = a + b exactly when |a| >= |b| or a == 0 (Dekker).
    __device__ static __forceinline__ void fast_two_sum(float a, float b, float &s, float &e)
    {
#pragma clang fp contract(off)
        s = a + b;
        e = b - (s - a);
    }
    __device__ static __forceinline__ T from(double v)
    {
#pragma clang fp contract(off)
        const float hi = static_cast<float>(v);
        return dw{hi, static_cast<float>(v - static_cast<double>(hi))};
    }
    __device__ static __forceinline__ double to(T v) { return static_cast<double>(v.hi) + static_cast<double>(v.lo); }
    __device__ static __forceinline__ T zero() { return dw{0.0f, 0.0f}; }
    __device__ static __forceinline__ T neg(T a) { return dw{-a.hi, -a.lo}; }
    // AccurateDWPlusDW.
    __device__ static __forceinline__ T add(T x, T y)
    {
#pragma clang fp contract(off)
        float sh, sl, th, tl;
        two_sum(x.hi, y.hi, sh, sl);
        two_sum(x.lo, y.lo, th, tl);
        sl = sl + th;
        float vh, vl;
        fast_two_sum(sh, sl, vh, vl);
        vl = vl + tl;
        float zh, zl;
        fast_two_sum(vh, vl, zh, zl);
        return dw{zh, zl};
    }
    __device__ static __forceinline__ T sub(T x, T y) { return add(x, neg(y)); }
    // DWTimesDW3.
    __device__ static __forceinline__ T mul(T x, T y)
    {
#pragma clang fp contract(off)
        const float ch = x.hi * y.hi;
        const float cl1 = __builtin_fmaf(x.hi, y.hi, -ch);
        const float tl0 = x.lo * y.lo;
        const float tl1 = __builtin_fmaf(x.hi, y.lo, tl0);
        const float cl2 = __builtin_fmaf(x.lo, y.hi, tl1);
        const float cl3 = cl1 + cl2;
        float zh, zl;
        fast_two_sum(ch, cl3, zh, zl);
        return dw{zh, zl};
    }
    __device__ static __forceinline__ T fma(T a, T b, T c) { return add(mul(a, b), c); }
    __device__ static __forceinline__ T scale2(T a, float b) { return dw{a.hi * b, a.lo * b}; }
};

}  // namespace prec
}  // namespace sfmm
