// FP64 per-term device functions shared by the exception kernel and the MD
// step's fused bonded kernel.  Each returns the term's energy and writes the
// forces on its particles; the caller accumulates them.
#pragma once

#include "common.h"

namespace ommhip {

// Exception pair (nonbonded.rs:995-1036): LJ and bare Coulomb with the pair's
// own parameters (lj_coulomb_exception, 1194-1198), minus the reciprocal-space
// erf correction with the particle charges (reciprocal_pair, 1275-1288) when
// a reciprocal method is active.  Force on i is fi, on j is -fi.
__device__ __forceinline__ double exception_xyz(const double pi[3], const double pj[3], double qq, double sig,
                                                double eps, double qi, double qj, const Box& box, int use_pbc,
                                                int has_pme, double alpha, double fi[3]) {
    double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
    if (use_pbc) min_image_exact(d, box.v);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = sqrt(r2);
    double e = 0.0, f = 0.0;
    if (eps != 0.0) {
        double sr2 = sig * sig / r2;
        double sr6 = sr2 * sr2 * sr2;
        double sr12 = sr6 * sr6;
        e += 4.0 * eps * (sr12 - sr6);
        f += 4.0 * eps * (-12.0 * sr12 + 6.0 * sr6) / r2;
    }
    if (qq != 0.0) {
        e += ONE_4PI_EPS0 * qq / r;
        f -= ONE_4PI_EPS0 * qq / (r2 * r);
    }
    if (has_pme) {
        double c = ONE_4PI_EPS0 * qi * qj;
        double corr, deriv;
        if (alpha * r < 1e-6) {
            corr = c * 2.0 * alpha / sqrt(PI) * (1.0 - alpha * alpha * r2 / 3.0);
            deriv = -c * 4.0 * alpha * alpha * alpha / (3.0 * sqrt(PI));
        } else {
            double erf_v = erf(alpha * r);
            corr = c * erf_v / r;
            deriv = c * (2.0 * alpha / sqrt(PI) * exp(-alpha * alpha * r2) - erf_v / r) / r2;
        }
        e -= corr;
        f -= deriv;
    }
    for (int a = 0; a < 3; ++a) fi[a] = f * d[a];
    return e;
}

__device__ __forceinline__ double exception_pair(int i, int j, double qq, double sig, double eps, double qi, double qj,
                                                 const double* __restrict__ pos, const Box& box, int use_pbc,
                                                 int has_pme, double alpha, double fi[3]) {
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    return exception_xyz(pi, pj, qq, sig, eps, qi, qj, box, use_pbc, has_pme, alpha, fi);
}

// HarmonicBondForce term without PBC (harmonic_bond.rs:113-140).  Force on i
// is fi, on j is -fi.
__device__ __forceinline__ double bond_xyz(const double pi[3], const double pj[3], double length, double k,
                                           double fi[3]) {
    double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double dr = r - length;
    double f = k * dr / r;
    for (int a = 0; a < 3; ++a) fi[a] = f * d[a];
    return 0.5 * k * dr * dr;
}

__device__ __forceinline__ double bond_pair(int i, int j, double length, double k, const double* __restrict__ pos,
                                            double fi[3]) {
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    return bond_xyz(pi, pj, length, k, fi);
}

// HarmonicAngleForce term without PBC (harmonic_angle.rs:118-167); j is the
// vertex.  Forces on i, j, k are fi, -(fi + fk), fk.
__device__ __forceinline__ double angle_xyz(const double pi[3], const double pj[3], const double pk[3], double theta0,
                                            double ka, double fi[3], double fk[3]) {
    double rji[3], rjk[3];
    for (int a = 0; a < 3; ++a) {
        rji[a] = pi[a] - pj[a];
        rjk[a] = pk[a] - pj[a];
    }
    double lji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    double ljk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    double c = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (lji * ljk);
    c = fmin(fmax(c, -1.0), 1.0);
    double dth = acos(c) - theta0;
    double s = fmax(sqrt(1.0 - c * c), 1e-8);
    double de = ka * dth;
    for (int a = 0; a < 3; ++a) {
        fi[a] = -de * (c * rji[a] / lji - rjk[a] / ljk) / (lji * s);
        fk[a] = -de * (c * rjk[a] / ljk - rji[a] / lji) / (ljk * s);
    }
    return 0.5 * ka * dth * dth;
}

__device__ __forceinline__ double angle_triple(int i, int j, int k, double theta0, double ka,
                                               const double* __restrict__ pos, double fi[3], double fk[3]) {
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    const double pk[3] = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
    return angle_xyz(pi, pj, pk, theta0, ka, fi, fk);
}

// Accumulating forms: add the forces to the force accumulator.
template <class A>
__device__ __forceinline__ double exception_term(int i, int j, double qq, double sig, double eps, double qi, double qj,
                                                 const double* __restrict__ pos, const Box& box, int use_pbc,
                                                 int has_pme, double alpha, A* force, unsigned* status) {
    double fi[3];
    const double e = exception_pair(i, j, qq, sig, eps, qi, qj, pos, box, use_pbc, has_pme, alpha, fi);
    for (int a = 0; a < 3; ++a) {
        acc_add(force + 3 * i + a, fi[a], FORCE_SCALE, status);
        acc_add(force + 3 * j + a, -fi[a], FORCE_SCALE, status);
    }
    return e;
}

template <class A>
__device__ __forceinline__ double bond_term(int i, int j, double length, double k, const double* __restrict__ pos,
                                            A* force, unsigned* status) {
    double fi[3];
    const double e = bond_pair(i, j, length, k, pos, fi);
    for (int a = 0; a < 3; ++a) {
        acc_add(force + 3 * i + a, fi[a], FORCE_SCALE, status);
        acc_add(force + 3 * j + a, -fi[a], FORCE_SCALE, status);
    }
    return e;
}

template <class A>
__device__ __forceinline__ double angle_term(int i, int j, int k, double theta0, double ka,
                                             const double* __restrict__ pos, A* force, unsigned* status) {
    double fi[3], fk[3];
    const double e = angle_triple(i, j, k, theta0, ka, pos, fi, fk);
    for (int a = 0; a < 3; ++a) {
        acc_add(force + 3 * i + a, fi[a], FORCE_SCALE, status);
        acc_add(force + 3 * k + a, fk[a], FORCE_SCALE, status);
        acc_add(force + 3 * j + a, -(fi[a] + fk[a]), FORCE_SCALE, status);
    }
    return e;
}

}  // namespace ommhip
