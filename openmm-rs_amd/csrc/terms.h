// FP64 per-term device functions shared by the exception kernel and the MD
// step's fused bonded kernel.  Each returns the term's energy and writes the
// forces on its particles; the caller accumulates them.
#pragma once

#include "common.h"

namespace ommhip {

// LJPME reciprocal-space dispersion of a pair (nonbonded.rs:1158-1191):
// E = C6 (1 - e^{-x} (1 + x + x^2/2)) / r^6 with x = alpha^2 r^2, and
// dE/dr / r; below x = 0.01 the power series of (1 - e^{-x}(1 + x + x^2/2))/x^3
// (terms t_k = t_{k-1} (-x)(k + 2)/((k + 3) k), t_0 = 1/6) avoids the
// cancellation.
__device__ __forceinline__ double disp_recip_pair(double c6, double r2, double alpha, double* deriv) {
    if (c6 == 0.0) {
        *deriv = 0.0;
        return 0.0;
    }
    const double x = alpha * alpha * r2;
    if (x < 0.01) {
        double term = 1.0 / 6.0, value = term, d = 0.0;
        for (int k = 1; k <= 12; ++k) {
            term *= -x * (k + 2) / (double)((k + 3) * k);
            value += term;
            d += x == 0.0 ? (k == 1 ? -1.0 / 8.0 : 0.0) : k * term / x;
        }
        const double a2 = alpha * alpha, a6 = a2 * a2 * a2;
        *deriv = 2.0 * c6 * a6 * a2 * d;
        return c6 * a6 * value;
    }
    const double ex = exp(-x), inv_r6 = 1.0 / (r2 * r2 * r2);
    const double poly = 1.0 + x + 0.5 * x * x;
    *deriv = -6.0 * c6 * inv_r6 / r2 * (1.0 - ex * (poly + x * x * x / 6.0));
    return c6 * inv_r6 * (1.0 - ex * poly);
}

// Exception pair (nonbonded.rs:995-1036): LJ and bare Coulomb with the pair's
// own parameters (lj_coulomb_exception, 1194-1198), minus the reciprocal-space
// erf correction with the particle charges (reciprocal_pair, 1275-1288) when
// a reciprocal method is active.  Force on i is fi, on j is -fi.
__device__ __forceinline__ double exception_xyz(const double pi[3], const double pj[3], double qq, double sig,
                                                double eps, double qi, double qj, const Box& box, int use_pbc,
                                                int has_pme, double alpha, double c6, double lj_alpha,
                                                double fi[3]) {
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
    if (c6 != 0.0) {
        // LJPME: the particles' reciprocal-space dispersion is added back.
        double deriv;
        e += disp_recip_pair(c6, r2, lj_alpha, &deriv);
        f += deriv;
    }
    for (int a = 0; a < 3; ++a) fi[a] = f * d[a];
    return e;
}

__device__ __forceinline__ double exception_pair(int i, int j, double qq, double sig, double eps, double qi, double qj,
                                                 const double* __restrict__ pos, const Box& box, int use_pbc,
                                                 int has_pme, double alpha, double c6, double lj_alpha,
                                                 double fi[3]) {
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    return exception_xyz(pi, pj, qq, sig, eps, qi, qj, box, use_pbc, has_pme, alpha, c6, lj_alpha, fi);
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
                                                 int has_pme, double alpha, double c6, double lj_alpha, A* force,
                                                 unsigned* status) {
    double fi[3];
    const double e = exception_pair(i, j, qq, sig, eps, qi, qj, pos, box, use_pbc, has_pme, alpha, c6, lj_alpha, fi);
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
