// Order-5 cardinal B-splines of smooth PME on the device: the weights and
// derivatives of one fractional coordinate and the stencil of a position.
#pragma once

#include "common.h"

namespace ommhip {

// Order-5 cardinal B-spline weights and derivatives (pme.rs:151-174).
template <class R>
__device__ __forceinline__ void spline5(R f, R w[5], R dw[5]) {
    w[0] = (R)1 - f;
    w[1] = f;
    w[2] = f * w[1] / (R)2;
    w[1] = ((f + (R)1) * w[0] + ((R)2 - f) * w[1]) / (R)2;
    w[0] *= ((R)1 - f) / (R)2;
    w[3] = f * w[2] / (R)3;
    w[2] = ((f + (R)1) * w[1] + ((R)3 - f) * w[2]) / (R)3;
    w[1] = ((f + (R)2) * w[0] + ((R)2 - f) * w[1]) / (R)3;
    w[0] *= ((R)1 - f) / (R)3;
    dw[0] = -w[0];
    dw[1] = w[0] - w[1];
    dw[2] = w[1] - w[2];
    dw[3] = w[2] - w[3];
    dw[4] = w[3];
    w[4] = f * w[3] / (R)4;
    w[3] = ((f + (R)1) * w[2] + ((R)4 - f) * w[3]) / (R)4;
    w[2] = ((f + (R)2) * w[1] + ((R)3 - f) * w[2]) / (R)4;
    w[1] = ((f + (R)3) * w[0] + ((R)2 - f) * w[1]) / (R)4;
    w[0] *= ((R)1 - f) / (R)4;
}

// pme.rs:201-218: base index and spline of the fractional grid coordinate.
template <class R>
__device__ __forceinline__ void stencil(const double* p, const Box& b, const int dims[3], int idx[3], R w[3][5],
                                        R dw[3][5]) {
    for (int a = 0; a < 3; ++a) {
        double f = p[0] * b.recip[a * 3] + p[1] * b.recip[a * 3 + 1] + p[2] * b.recip[a * 3 + 2];
        double c = (f - floor(f)) * (double)dims[a];
        double fl = floor(c);
        idx[a] = ((int)fl) % dims[a];
        spline5<R>((R)(c - fl), w[a], dw[a]);
    }
}

}  // namespace ommhip
