// Arithmetic helpers shared by the FP32 and FP64 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ommhip {

template <class R>
struct Math;

// FP32 pair arithmetic.  erfc(x) = erfcx(x) * exp(-x^2) with erfcx fitted
// (this project's own weighted Chebyshev fit, converted to monomials) as a
// degree-8 polynomial in t = (x - 2.5)/(x + 2.5) on x in [0, 6]: relative error
// 1.2e-8 in exact arithmetic, 1.6e-7 evaluated in FP32.  exp(-x^2) reuses the
// hardware exp2; reciprocal and reciprocal square root are the hardware's
// (about 1 ulp), well inside the FP32 rounding of the force sums.
template <>
struct Math<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ void erfc_exp(float x, float& erfc_v, float& ex) {
        float t = (x - 2.5f) * rcp(x + 2.5f);
        t = fminf(t, 0.41176470f);
        float p = 0.0003302688419353217f;
        p = fmaf(p, t, 0.0008612216333858669f);
        p = fmaf(p, t, -0.0026842965744435787f);
        p = fmaf(p, t, -0.003998667001724243f);
        p = fmaf(p, t, 0.039928220212459564f);
        p = fmaf(p, t, -0.12503232061862946f);
        p = fmaf(p, t, 0.2517129182815552f);
        p = fmaf(p, t, -0.3717367649078369f);
        p = fmaf(p, t, 0.21080636978149414f);
        ex = __builtin_amdgcn_exp2f(-(x * x) * 1.4426950408889634f);
        erfc_v = p * ex;
    }
    static __device__ __forceinline__ float round(float x) { return rintf(x); }
    static __device__ __forceinline__ float max(float a, float b) { return fmaxf(a, b); }
    // e^{-x} through the hardware exp2.
    static __device__ __forceinline__ float exp_neg(float x) { return __builtin_amdgcn_exp2f(-x * 1.4426950408889634f); }
};

template <>
struct Math<double> {
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / sqrt(x); }
    static __device__ __forceinline__ void erfc_exp(double x, double& erfc_v, double& ex) {
        erfc_v = erfc(x);
        ex = exp(-x * x);
    }
    static __device__ __forceinline__ double round(double x) { return rint(x); }
    static __device__ __forceinline__ double max(double a, double b) { return fmax(a, b); }
    static __device__ __forceinline__ double exp_neg(double x) { return exp(-x); }
};

}  // namespace ommhip
