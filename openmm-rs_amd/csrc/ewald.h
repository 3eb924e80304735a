// Plain Ewald reciprocal space (ewald.rs:22-122) for one rectangular periodic
// system: the structure factor S(k) = sum_i q_i e^{i k.r_i} over the
// half-space of wave vectors k = 2 pi (x/Lx, y/Ly, z/Lz) with |x| < kx,
// |y| < ky, |z| < kz, the energy sum_k I(k) |S(k)|^2 and the forces
// 2 I(k) Im(conj(S(k)) q_i e^{i k.r_i}) k with
// I(k) = ONE_4PI_EPS0 4 pi / V e^{-k^2 / 4 alpha^2} / k^2.
//
// The device form: a table of per-atom phases e^{2 pi i m x_a / L_a}
// (m < kmax, each axis; FP64 sincospi of the exact argument), one workgroup
// per wave vector reducing S(k) over the atoms in a fixed order, and one
// thread per atom summing its force over the wave vectors (the same
// products, so no phase is stored per (k, atom)).  Everything is FP64 in
// both precisions: Ewald is the cross-check method beside PME, with
// O(N K) work.
#pragma once

#include "common.h"

namespace ommhip {

// ewald.rs:22-46: alpha and the per-axis wave-vector limits (odd).
void ewald_parameters(const double widths[3], double cutoff, double tolerance, double* alpha, int limits[3]);

struct EwaldDev {
    int n = 0, kmax = 0, nk = 0;
    int limits[3] = {};
    DevBuf<int4> kvec;     // [nk] (x, y, z, 0) in the reference's loop order
    DevBuf<double2> phase; // [3][kmax][n] e^{2 pi i m x_a / L_a}
    DevBuf<double2> sfac;  // [nk] I(k) S(k)
};

void ewald_setup(EwaldDev& ew, int n, const int limits[3], hipStream_t s);

// Adds the reciprocal forces (and, with an energy buffer, the energy into
// slot E_RECIP of system 0) for the box of sys_dev[0] (its diagonal).
void ewald_run(EwaldDev& ew, const double* pos, const double* charge, const SysInfo* sys_dev, double alpha,
               long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
               bool mixed, hipStream_t s);

}  // namespace ommhip
