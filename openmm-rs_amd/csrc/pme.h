// Smooth PME reciprocal space (order-5 B-splines), pme.rs:151-425.
#pragma once

#include "pairs.h"

namespace ommhip {

struct PmeDev {
    int nx = 0, ny = 0, nz = 0, nzh = 0;
    double alpha = 0.0;
    std::vector<int> radix[3];
    DevBuf<long long> qgrid_fx;  // mixed: fixed-point charge grid [nx][ny][nz]
    DevBuf<double> qgrid_d;      // plain: FP64 charge grid
    DevBuf<float> cgrid_f;       // complex half-spectrum [nx][ny][nzh] (interleaved re, im)
    DevBuf<double> cgrid_d;
    DevBuf<float> phi_f;         // real-space potential [nx][ny][nz]
    DevBuf<double> phi_d;
    DevBuf<float> roots_f;       // per axis n roots e^{-2 pi i m/n}, interleaved (x, then y, then z)
    DevBuf<double> roots_d;
    DevBuf<float> moduli_f;      // B-spline moduli, x then y then z
    DevBuf<double> moduli_d;
    DevBuf<int> radix_dev;       // [3][8] radices per axis, 0-terminated
};

// Grid dimensions must each be >= 5 (pme.rs:222-225) and at most 1024.
void pme_setup(PmeDev& pm, double alpha, const int dims[3], bool mixed, hipStream_t s);

// Spread, FFT, convolution (+energy into slot E_RECIP), inverse FFT, gather
// (forces added).  Atoms with zero charge are skipped as in the reference.
enum : int { PME_SPREAD = 1, PME_FFT = 2, PME_GATHER = 4, PME_ALL = 7 };
// Arithmetic of one PME evaluation.
//   MESH_F32      : FP32 splines/FFT/gather, fixed-point grid and forces (mixed
//                   form, force-only MD steps);
//   MESH_F64_FIXED: FP64 splines/FFT/gather on the same fixed-point grid with
//                   fixed-point forces (mixed form when the energy is wanted:
//                   the mesh energy is ~1e4-1e5 kJ/mol and must carry ~1e-9
//                   relative error);
//   MESH_F64      : FP64 throughout with FP64 atomics (plain form).
enum MeshMode : int { MESH_F32 = 0, MESH_F64_FIXED = 1, MESH_F64 = 2 };
void pme_run(PmeDev& pm, int mode, int n, const double* pos, const double* charge, const Box* box_dev,
             long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
             hipStream_t s, int phases = PME_ALL);

// B-spline moduli |sum_i M5(i+1) e^{2 pi i k i/n}|^2 with near-zero values
// replaced by the mean of their neighbours (pme.rs:239-259).
std::vector<double> bspline_moduli(int n);

}  // namespace ommhip
