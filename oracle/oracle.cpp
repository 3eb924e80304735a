// oracle.cpp — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// FP64 restatement of crates/openmm-core/src/forces/{nonbonded,pme,neighbors,
// ewald,harmonic_bond,harmonic_angle}.rs, integrators/langevin.rs,
// forces/barostat.rs and the Context helpers they use.  Single-threaded like
// the reference (no rayon/thread use in its crates).  Compile with
// -ffp-contract=off.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <set>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
// nonbonded.rs:16
constexpr double ONE_4PI_EPS0 = 138.93545764438198;
// langevin.rs:12
const double BOLTZMANN = 1.380649e-23 * 6.02214076e23 / 1000.0;
// barostat.rs:13
constexpr double BAR_NM3_TO_KJ_PER_MOL = 0.0602214076;

using cd = std::complex<double>;

struct Particle {
    double charge, sigma, epsilon;
};
struct Exception {
    int p1, p2;
    double charge_prod, sigma, epsilon;
};

// harmonic_bond.rs:152-159 — reduced-triclinic minimum image, c then b then a.
inline void min_image(double d[3], const double* bv) {
    for (int dd = 2; dd >= 0; --dd) {
        double shift = std::floor(d[dd] / bv[dd * 3 + dd] + 0.5);
        for (int axis = 0; axis <= dd; ++axis) {
            double t = shift * bv[dd * 3 + axis];
            d[axis] -= t;
        }
    }
}

// pme.rs:176-199
double reciprocal_box(const double* bv, double recip[9]) {
    auto cross = [](const double* a, const double* b, double* out) {
        out[0] = a[1] * b[2] - a[2] * b[1];
        out[1] = a[2] * b[0] - a[0] * b[2];
        out[2] = a[0] * b[1] - a[1] * b[0];
    };
    const double* a = bv;
    const double* b = bv + 3;
    const double* c = bv + 6;
    double bc[3], ca[3], ab[3];
    cross(b, c, bc);
    cross(c, a, ca);
    cross(a, b, ab);
    double volume = a[0] * bc[0];
    volume += a[1] * bc[1];
    volume += a[2] * bc[2];
    for (int i = 0; i < 3; ++i) {
        recip[i] = bc[i] / volume;
        recip[3 + i] = ca[i] / volume;
        recip[6 + i] = ab[i] / volume;
    }
    return volume;
}

// pme.rs:151-174 — order-5 cardinal B-spline weights and derivatives.
void spline(double f, double w[5], double dw[5]) {
    for (int i = 0; i < 5; ++i) {
        w[i] = 0.0;
        dw[i] = 0.0;
    }
    w[0] = 1.0 - f;
    w[1] = f;
    for (int order = 3; order <= 5; ++order) {
        if (order == 5) {
            dw[0] = -w[0];
            for (int i = 1; i < 5; ++i) dw[i] = w[i - 1] - w[i];
        }
        double den = (double)(order - 1);
        w[order - 1] = f * w[order - 2] / den;
        for (int i = order - 2; i >= 1; --i) {
            double a = (f + (double)(order - 1 - i)) * w[i - 1];
            double b = ((double)(i + 1) - f) * w[i];
            w[i] = (a + b) / den;
        }
        w[0] *= (1.0 - f) / den;
    }
}

// Mixed-radix complex FFT (rustfft conventions: forward e^{-2 pi i kn/N},
// inverse e^{+2 pi i kn/N}, no normalisation).
struct FftPlan {
    int n = 0;
    std::vector<cd> roots;  // roots[k] = e^{-2 pi i k/n}
    explicit FftPlan(int n_) : n(n_), roots(n_) {
        for (int k = 0; k < n; ++k) {
            double t = -2.0 * kPi * (double)k / (double)n;
            roots[k] = cd(std::cos(t), std::sin(t));
        }
    }
    // out[k] = sum_j in[j*stride] * W^{jk}, W = e^{sign 2 pi i / m}, m = len.
    void rec(const cd* in, cd* out, int m, int stride, bool inverse, cd* tmp) const {
        if (m == 1) {
            out[0] = in[0];
            return;
        }
        int p = 2;
        while (p * p <= m && m % p) ++p;
        if (m % p) p = m;  // prime
        int q = m / p;
        int step = n / m;  // root index scale for this sub-length
        if (q == 1) {
            for (int k = 0; k < m; ++k) {
                cd s(0.0, 0.0);
                for (int j = 0; j < m; ++j) {
                    int idx = (int)(((long long)j * k) % m) * step;
                    cd r = inverse ? std::conj(roots[idx]) : roots[idx];
                    s += in[(long long)j * stride] * r;
                }
                out[k] = s;
            }
            return;
        }
        // p sub-transforms of length q on decimated inputs.
        for (int r = 0; r < p; ++r) rec(in + (long long)r * stride, out + r * q, q, stride * p, inverse, tmp + m);
        for (int k = 0; k < m; ++k) {
            cd s(0.0, 0.0);
            int kq = k % q;
            for (int r = 0; r < p; ++r) {
                int idx = (int)(((long long)r * k) % m) * step;
                cd w = inverse ? std::conj(roots[idx]) : roots[idx];
                s += out[r * q + kq] * w;
            }
            tmp[k] = s;
        }
        for (int k = 0; k < m; ++k) out[k] = tmp[k];
    }
    void run(cd* data, bool inverse) const {
        std::vector<cd> in(data, data + n), scratch(8 * n + 64);
        rec(in.data(), data, n, 1, inverse, scratch.data());
    }
};

// pme.rs:221-269 (moduli part).
std::vector<double> bspline_moduli(int n) {
    double w[5], dw[5];
    spline(0.0, w, dw);
    std::vector<double> values(n);
    for (int k = 0; k < n; ++k) {
        cd sum(0.0, 0.0);
        for (int i = 0; i < 5; ++i) {
            double t = 2.0 * kPi * (double)k * (double)i / (double)n;
            sum += cd(w[i] * std::cos(t), w[i] * std::sin(t));
        }
        values[k] = sum.real() * sum.real() + sum.imag() * sum.imag();
    }
    for (int k = 0; k < n; ++k)
        if (values[k] < 1e-7) values[k] = 0.5 * (values[(k + n - 1) % n] + values[(k + 1) % n]);
    return values;
}

// pme.rs:201-218
void stencil(const double* position, const double* recip, const int* dims, int idx[3],
             double w[3][5], double dw[3][5]) {
    for (int axis = 0; axis < 3; ++axis) {
        double fractional = position[0] * recip[axis * 3 + 0];
        fractional += position[1] * recip[axis * 3 + 1];
        fractional += position[2] * recip[axis * 3 + 2];
        double coordinate = (fractional - std::floor(fractional)) * (double)dims[axis];
        idx[axis] = (int)((long long)std::floor(coordinate) % dims[axis]);
        spline(coordinate - std::floor(coordinate), w[axis], dw[axis]);
    }
}

struct PmeWorkspace {
    int dims[3];
    std::vector<cd> grid;
    std::vector<double> moduli[3];
    FftPlan* plans[3];
    explicit PmeWorkspace(const int* d) {
        for (int a = 0; a < 3; ++a) dims[a] = d[a];
        grid.assign((size_t)d[0] * d[1] * d[2], cd(0.0, 0.0));
        for (int a = 0; a < 3; ++a) {
            moduli[a] = bspline_moduli(d[a]);
            plans[a] = new FftPlan(d[a]);
        }
    }
    ~PmeWorkspace() {
        for (auto* p : plans) delete p;
    }
    // pme.rs:271-303
    void transform(bool inverse) {
        int nx = dims[0], ny = dims[1], nz = dims[2];
        for (size_t row = 0; row < (size_t)nx * ny; ++row) plans[2]->run(&grid[row * nz], inverse);
        std::vector<cd> line(std::max(nx, ny));
        for (int x = 0; x < nx; ++x)
            for (int z = 0; z < nz; ++z) {
                for (int y = 0; y < ny; ++y) line[y] = grid[((size_t)x * ny + y) * nz + z];
                plans[1]->run(line.data(), inverse);
                for (int y = 0; y < ny; ++y) grid[((size_t)x * ny + y) * nz + z] = line[y];
            }
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                for (int x = 0; x < nx; ++x) line[x] = grid[((size_t)x * ny + y) * nz + z];
                plans[0]->run(line.data(), inverse);
                for (int x = 0; x < nx; ++x) grid[((size_t)x * ny + y) * nz + z] = line[x];
            }
    }
    // pme.rs:305-375
    double potential(const double* pos, const double* amp, int n, const double* bv, double alpha,
                     bool dispersion) {
        double recip[9];
        double volume = reciprocal_box(bv, recip);
        int nx = dims[0], ny = dims[1], nz = dims[2];
        std::fill(grid.begin(), grid.end(), cd(0.0, 0.0));
        for (int a = 0; a < n; ++a) {
            double charge = amp[a];
            if (charge == 0.0) continue;
            int idx[3];
            double w[3][5], dw[3][5];
            stencil(pos + 3 * a, recip, dims, idx, w, dw);
            for (int ix = 0; ix < 5; ++ix)
                for (int iy = 0; iy < 5; ++iy)
                    for (int iz = 0; iz < 5; ++iz) {
                        size_t index = ((size_t)((idx[0] + ix) % nx) * ny + (idx[1] + iy) % ny) * nz +
                                       (idx[2] + iz) % nz;
                        grid[index] += cd(charge * w[0][ix] * w[1][iy] * w[2][iz], 0.0);
                    }
        }
        transform(false);
        auto frequency = [](int k, int nn) {
            return k < (nn + 1) / 2 ? (double)k : (double)k - (double)nn;
        };
        double energy = 0.0;
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                for (int z = 0; z < nz; ++z) {
                    size_t index = ((size_t)x * ny + y) * nz + z;
                    if (index == 0 && !dispersion) {
                        grid[index] = cd(0.0, 0.0);
                        continue;
                    }
                    double modes[3] = {frequency(x, nx), frequency(y, ny), frequency(z, nz)};
                    double k2 = 0.0;
                    for (int axis = 0; axis < 3; ++axis) {
                        double s = modes[0] * recip[0 * 3 + axis];
                        s += modes[1] * recip[1 * 3 + axis];
                        s += modes[2] * recip[2 * 3 + axis];
                        k2 += s * s;
                    }
                    double mod = moduli[0][x] * moduli[1][y] * moduli[2][z];
                    double influence;
                    if (dispersion) {
                        double k = std::sqrt(k2);
                        double b = kPi * k / alpha;
                        influence = -kPi * std::sqrt(kPi) / (3.0 * volume * mod) *
                                    (2.0 * std::pow(kPi, 3) * std::sqrt(kPi) * std::erfc(b) * k * k2 +
                                     std::exp(-b * b) *
                                         (std::pow(alpha, 3) - 2.0 * alpha * kPi * kPi * k2));
                    } else {
                        influence = ONE_4PI_EPS0 * std::exp(-kPi * kPi * k2 / (alpha * alpha)) /
                                    (kPi * volume * k2 * mod);
                    }
                    cd g = grid[index];
                    energy += 0.5 * influence * (g.real() * g.real() + g.imag() * g.imag());
                    grid[index] = g * influence;
                }
        transform(true);
        return energy;
    }
    // pme.rs:377-425
    double evaluate(const double* pos, const double* amp, int n, const double* bv, double alpha,
                    bool dispersion, double* forces) {
        double energy = potential(pos, amp, n, bv, alpha, dispersion);
        double recip[9];
        reciprocal_box(bv, recip);
        int nx = dims[0], ny = dims[1], nz = dims[2];
        if (!forces) return energy;
        for (int a = 0; a < n; ++a) {
            double charge = amp[a];
            if (charge == 0.0) continue;
            int idx[3];
            double w[3][5], dw[3][5];
            stencil(pos + 3 * a, recip, dims, idx, w, dw);
            double g[3] = {0.0, 0.0, 0.0};
            for (int ix = 0; ix < 5; ++ix)
                for (int iy = 0; iy < 5; ++iy)
                    for (int iz = 0; iz < 5; ++iz) {
                        size_t index = ((size_t)((idx[0] + ix) % nx) * ny + (idx[1] + iy) % ny) * nz +
                                       (idx[2] + iz) % nz;
                        double phi = grid[index].real();
                        g[0] += phi * dw[0][ix] * w[1][iy] * w[2][iz];
                        g[1] += phi * w[0][ix] * dw[1][iy] * w[2][iz];
                        g[2] += phi * w[0][ix] * w[1][iy] * dw[2][iz];
                    }
            for (int axis = 0; axis < 3; ++axis) {
                double s = g[0] * (double)dims[0] * recip[0 * 3 + axis];
                s += g[1] * (double)dims[1] * recip[1 * 3 + axis];
                s += g[2] * (double)dims[2] * recip[2 * 3 + axis];
                forces[3 * a + axis] -= charge * s;
            }
        }
        return energy;
    }
};

// nonbonded.rs:1203-1214
inline void lj_pair(double sigma, double epsilon, double r2, double& e, double& f) {
    if (epsilon == 0.0) {
        e = 0.0;
        f = 0.0;
        return;
    }
    double sr2 = (sigma * sigma) / r2;
    double sr6 = sr2 * sr2 * sr2;
    double sr12 = sr6 * sr6;
    e = 4.0 * epsilon * (sr12 - sr6);
    f = 4.0 * epsilon * (-12.0 * sr12 + 6.0 * sr6) / r2;
}

// nonbonded.rs:1219-1226
inline void coulomb_pair(double qq, double r, double r2, double& e, double& f) {
    if (qq == 0.0) {
        e = 0.0;
        f = 0.0;
        return;
    }
    e = ONE_4PI_EPS0 * qq / r;
    f = -ONE_4PI_EPS0 * qq / (r2 * r);
}

// nonbonded.rs:1233-1250
inline void reaction_field(double qq, double r, double r2, double rc, double eps_rf, double& e,
                           double& f) {
    if (qq == 0.0) {
        e = 0.0;
        f = 0.0;
        return;
    }
    double rc3 = rc * rc * rc;
    double k_rf, c_rf;
    if (std::isinf(eps_rf)) {
        k_rf = 0.5 / rc3;
        c_rf = 1.5 / rc;
    } else {
        k_rf = (eps_rf - 1.0) / (2.0 * eps_rf + 1.0) / rc3;
        c_rf = (3.0 * eps_rf) / (2.0 * eps_rf + 1.0) / rc;
    }
    e = ONE_4PI_EPS0 * qq * (1.0 / r + k_rf * r2 - c_rf);
    f = ONE_4PI_EPS0 * qq * (-1.0 / (r2 * r) + 2.0 * k_rf);
}

// nonbonded.rs:1263-1273
inline void screened_coulomb(double qq, double r, double r2, double alpha, double& e, double& f) {
    if (qq == 0.0) {
        e = 0.0;
        f = 0.0;
        return;
    }
    double erfc_v = std::erfc(alpha * r);
    double coefficient = ONE_4PI_EPS0 * qq;
    e = coefficient * erfc_v / r;
    f = -coefficient * (erfc_v / r + 2.0 * alpha / std::sqrt(kPi) * std::exp(-alpha * alpha * r2)) / r2;
}

// nonbonded.rs:1275-1288
inline void reciprocal_pair(double qq, double r, double r2, double alpha, double& e, double& f) {
    double coefficient = ONE_4PI_EPS0 * qq;
    if (alpha * r < 1e-6) {
        e = coefficient * 2.0 * alpha / std::sqrt(kPi) * (1.0 - alpha * alpha * r2 / 3.0);
        f = -coefficient * 4.0 * std::pow(alpha, 3) / (3.0 * std::sqrt(kPi));
        return;
    }
    double erf_v = std::erf(alpha * r);
    e = coefficient * erf_v / r;
    f = coefficient * (2.0 * alpha / std::sqrt(kPi) * std::exp(-alpha * alpha * r2) - erf_v / r) / r2;
}

// nonbonded.rs:1158-1191
inline void dispersion_reciprocal_pair(double c6, double r2, double alpha, double& e, double& f) {
    if (c6 == 0.0) {
        e = 0.0;
        f = 0.0;
        return;
    }
    double x = alpha * alpha * r2;
    if (x < 0.01) {
        double term = 1.0 / 6.0;
        double value = term;
        double derivative = 0.0;
        for (int k = 1; k <= 12; ++k) {
            term *= -x * (double)(k + 2) / (double)((k + 3) * k);
            value += term;
            derivative += (x == 0.0) ? (k == 1 ? -1.0 / 8.0 : 0.0) : (double)k * term / x;
        }
        e = c6 * std::pow(alpha, 6) * value;
        f = 2.0 * c6 * std::pow(alpha, 8) * derivative;
        return;
    }
    double ex = std::exp(-x);
    e = c6 / std::pow(r2, 3) * (1.0 - ex * (1.0 + x + 0.5 * x * x));
    f = -6.0 * c6 / std::pow(r2, 4) * (1.0 - ex * (1.0 + x + 0.5 * x * x + std::pow(x, 3) / 6.0));
}

// nonbonded.rs:1290-1336 — adaptive Simpson.
template <class F>
double simpson_recurse(const F& f, double a, double b, double fa, double fm, double fb,
                       double estimate, double tolerance, int depth) {
    double midpoint = 0.5 * (a + b);
    double fl = f(0.5 * (a + midpoint));
    double fr = f(0.5 * (midpoint + b));
    double left = (midpoint - a) / 6.0 * (fa + 4.0 * fl + fm);
    double right = (b - midpoint) / 6.0 * (fm + 4.0 * fr + fb);
    double difference = left + right - estimate;
    if (depth == 0 || std::fabs(difference) <= 15.0 * tolerance) return left + right + difference / 15.0;
    return simpson_recurse(f, a, midpoint, fa, fl, fm, left, 0.5 * tolerance, depth - 1) +
           simpson_recurse(f, midpoint, b, fm, fr, fb, right, 0.5 * tolerance, depth - 1);
}
template <class F>
double integrate(const F& f, double a, double b) {
    double v0 = f(a), v1 = f(0.5 * (a + b)), v2 = f(b);
    double estimate = (b - a) / 6.0 * (v0 + 4.0 * v1 + v2);
    return simpson_recurse(f, a, b, v0, v1, v2, estimate, 1e-11 * std::max(std::fabs(estimate), 1e-12), 20);
}

struct Desc {
    std::vector<Particle> particles;
    std::vector<Exception> exceptions;
    int method;
    double cutoff, rf_dielectric, tolerance, pme_alpha;
    int pme_grid[3];
    bool dispersion_correction, exceptions_pbc, switching;
    double switching_distance;
    explicit Desc(const omm_nb_desc* d) {
        particles.resize(d->num_particles);
        for (int i = 0; i < d->num_particles; ++i)
            particles[i] = {d->charge[i], d->sigma[i], d->epsilon[i]};
        exceptions.resize(d->num_exceptions);
        for (int i = 0; i < d->num_exceptions; ++i)
            exceptions[i] = {d->exception_particle1[i], d->exception_particle2[i],
                             d->exception_charge_prod[i], d->exception_sigma[i],
                             d->exception_epsilon[i]};
        method = d->method;
        cutoff = d->cutoff;
        rf_dielectric = d->reaction_field_dielectric;
        tolerance = d->ewald_error_tolerance;
        pme_alpha = d->pme_alpha;
        for (int a = 0; a < 3; ++a) pme_grid[a] = d->pme_grid[a];
        dispersion_correction = d->use_dispersion_correction != 0;
        exceptions_pbc = d->exceptions_use_pbc != 0;
        switching = d->use_switching_function != 0;
        switching_distance = d->switching_distance;
        apply_offsets(d);
    }
    // nonbonded.rs:395-430: parameter offsets at the global parameters'
    // values (here their defaults), walked last-first with the first
    // (parameter, target) seen winning.
    void apply_offsets(const omm_nb_desc* d) {
        std::set<std::pair<int, int>> seen;
        for (int k = d->num_particle_offsets - 1; k >= 0; --k) {
            const int prm = d->particle_offset_parameter[k], idx = d->particle_offset_index[k];
            if (!seen.insert({prm, idx}).second) continue;
            const double v = d->global_parameter_defaults[prm];
            const double* sc = d->particle_offset_scales + 3 * k;
            particles[idx].charge += v * sc[0];
            particles[idx].sigma += v * sc[1];
            particles[idx].epsilon += v * sc[2];
        }
        seen.clear();
        for (int k = d->num_exception_offsets - 1; k >= 0; --k) {
            const int prm = d->exception_offset_parameter[k], idx = d->exception_offset_index[k];
            if (!seen.insert({prm, idx}).second) continue;
            const double v = d->global_parameter_defaults[prm];
            const double* sc = d->exception_offset_scales + 3 * k;
            exceptions[idx].charge_prod += v * sc[0];
            exceptions[idx].sigma += v * sc[1];
            exceptions[idx].epsilon += v * sc[2];
        }
    }
    bool periodic() const {
        return method == OMM_NB_CUTOFF_PERIODIC || method == OMM_NB_EWALD || method == OMM_NB_PME ||
               method == OMM_NB_LJPME;
    }
};

// nonbonded.rs:660-719
double dispersion_coefficient(const Desc& d) {
    double moments[13] = {0}, diagonal[13] = {0};
    for (const auto& p : d.particles) {
        double power = 1.0;
        for (int k = 0; k <= 12; ++k) {
            moments[k] += std::sqrt(p.epsilon) * power;
            diagonal[k] += p.epsilon * power;
            power *= p.sigma;
        }
    }
    auto pair_moment = [&](int power) {
        double coefficient = 1.0, sum = 0.0;
        for (int k = 0; k <= power; ++k) {
            sum += coefficient * moments[k] * moments[power - k];
            coefficient *= (double)(power - k) / (double)(k + 1);
        }
        return sum / std::pow(2.0, power) + diagonal[power];
    };
    auto integral = [&](int power) {
        double tail = std::pow(d.cutoff, 3 - power) / (double)(power - 3);
        if (!d.switching) return tail;
        double sw = d.switching_distance, rc = d.cutoff;
        auto f = [&](double r) {
            double t = (r - sw) / (rc - sw);
            return std::pow(t, 3) * (10.0 + t * (-15.0 + 6.0 * t)) * std::pow(r, 2 - power);
        };
        return tail + integrate(f, sw, rc);
    };
    double n = (double)d.particles.size();
    return 8.0 * kPi * n / (n + 1.0) * (pair_moment(12) * integral(12) - pair_moment(6) * integral(6));
}

// neighbors.rs:23-181 — cell-list candidate enumeration (i<j).
template <class Visit>
void cell_visit(const double* pos, int n, double cutoff, const double* bv, Visit&& pair) {
    if (n < 2) return;
    long long dims[3];
    double origin[3] = {0, 0, 0};
    double recip[9];
    bool periodic = bv != nullptr;
    double inverse_cutoff = 1.0 / cutoff;
    if (periodic) {
        reciprocal_box(bv, recip);
        for (int a = 0; a < 3; ++a) {
            double norm = std::sqrt(recip[a * 3] * recip[a * 3] + recip[a * 3 + 1] * recip[a * 3 + 1] +
                                    recip[a * 3 + 2] * recip[a * 3 + 2]);
            double v = std::floor(1.0 / (cutoff * norm));
            v = std::min(std::max(v, 1.0), 1e9);
            dims[a] = (long long)v;
        }
    } else {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], pos[3 * i + a]);
                mx[a] = std::max(mx[a], pos[3 * i + a]);
            }
        for (int a = 0; a < 3; ++a) {
            dims[a] = (long long)std::floor((mx[a] - mn[a]) * inverse_cutoff) + 1;
            origin[a] = mn[a];
        }
    }
    long long full_span = periodic ? 3 : 2;
    if (dims[0] <= full_span && dims[1] <= full_span && dims[2] <= full_span) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) pair(i, j);
        return;
    }
    auto key = [&](const double* p, long long k[3]) {
        for (int a = 0; a < 3; ++a) {
            if (periodic) {
                double f = p[0] * recip[a * 3 + 0];
                f += p[1] * recip[a * 3 + 1];
                f += p[2] * recip[a * 3 + 2];
                long long v = (long long)std::floor((f - std::floor(f)) * (double)dims[a]);
                k[a] = std::min(v, dims[a] - 1);
            } else {
                k[a] = (long long)std::floor((p[a] - origin[a]) * inverse_cutoff);
            }
        }
    };
    auto offset = [&](const long long k[3]) { return (size_t)((k[0] * dims[1] + k[1]) * dims[2] + k[2]); };
    const uint32_t EMPTY = 0xffffffffu;
    bool dense = false;
    size_t dense_size = 1;
    {
        bool ok = true;
        for (int a = 0; a < 3 && ok; ++a) {
            if (dense_size > (size_t)-1 / (size_t)dims[a]) ok = false;
            else dense_size *= (size_t)dims[a];
        }
        dense = ok && dense_size <= (size_t)8 * n;
    }
    std::vector<uint32_t> heads, next(n, EMPTY);
    struct KeyHash {
        size_t operator()(const std::array<long long, 3>& k) const {
            return std::hash<long long>()(k[0] * 73856093LL ^ k[1] * 19349663LL ^ k[2] * 83492791LL);
        }
    };
    std::unordered_map<std::array<long long, 3>, uint32_t, KeyHash> sparse;
    if (dense) heads.assign(dense_size, EMPTY);
    for (int i = 0; i < n; ++i) {
        long long k[3];
        key(pos + 3 * i, k);
        uint32_t previous;
        if (dense) {
            size_t idx = offset(k);
            previous = heads[idx];
            heads[idx] = (uint32_t)i;
        } else {
            std::array<long long, 3> kk{k[0], k[1], k[2]};
            auto it = sparse.find(kk);
            if (it == sparse.end()) {
                previous = EMPTY;
                sparse.emplace(kk, (uint32_t)i);
            } else {
                previous = it->second;
                it->second = (uint32_t)i;
            }
        }
        next[i] = previous;
    }
    static const long long off1[] = {0}, off2[] = {0, 1}, off3[] = {-1, 0, 1};
    auto offsets = [&](int a, int& count) -> const long long* {
        if (periodic && dims[a] == 1) {
            count = 1;
            return off1;
        }
        if (periodic && dims[a] == 2) {
            count = 2;
            return off2;
        }
        count = 3;
        return off3;
    };
    for (int i = 0; i < n; ++i) {
        long long c[3];
        key(pos + 3 * i, c);
        int cx, cy, cz;
        const long long* ox = offsets(0, cx);
        const long long* oy = offsets(1, cy);
        const long long* oz = offsets(2, cz);
        for (int a = 0; a < cx; ++a)
            for (int b = 0; b < cy; ++b)
                for (int e = 0; e < cz; ++e) {
                    long long k[3] = {c[0] + ox[a], c[1] + oy[b], c[2] + oz[e]};
                    if (periodic) {
                        for (int ax = 0; ax < 3; ++ax) {
                            long long v = k[ax] % dims[ax];
                            if (v < 0) v += dims[ax];
                            k[ax] = v;
                        }
                    } else if (k[0] < 0 || k[0] >= dims[0] || k[1] < 0 || k[1] >= dims[1] || k[2] < 0 ||
                               k[2] >= dims[2]) {
                        continue;
                    }
                    uint32_t j;
                    if (dense) {
                        j = heads[offset(k)];
                    } else {
                        auto it = sparse.find({k[0], k[1], k[2]});
                        j = it == sparse.end() ? EMPTY : it->second;
                    }
                    while (j != EMPTY) {
                        if ((int)j > i) pair(i, (int)j);
                        j = next[j];
                    }
                }
    }
}

struct PairKeyHash {
    size_t operator()(uint64_t k) const { return (size_t)(k * 0x9E3779B97F4A7C15ull); }
};

uint64_t pair_key(int32_t i, int32_t j) {
    uint64_t z = ((uint64_t)(uint32_t)i << 32) | (uint64_t)(uint32_t)j;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void pme_params(double tolerance, double cutoff, const double* bv, double factor, double* alpha,
                int32_t* grid) {
    *alpha = std::sqrt(-std::log(2.0 * tolerance)) / cutoff;
    for (int a = 0; a < 3; ++a) {
        double v = std::ceil(factor * *alpha * bv[a * 3 + a] / (3.0 * std::pow(tolerance, 0.2)));
        grid[a] = (int32_t)std::max(v, 6.0);
    }
}

}  // namespace

extern "C" {

uint64_t orc_pair_key(int32_t i, int32_t j) { return pair_key(i, j); }

void orc_effective_pme_parameters(double tolerance, double cutoff, const double box[9], double* alpha,
                                  int32_t grid[3]) {
    pme_params(tolerance, cutoff, box, 2.0, alpha, grid);
}

void orc_effective_ljpme_parameters(double tolerance, double cutoff, const double box[9], double* alpha,
                                    int32_t grid[3]) {
    pme_params(tolerance, cutoff, box, 1.0, alpha, grid);
}

// ewald.rs:22-46
void orc_ewald_parameters(const double widths[3], double cutoff, double tolerance, double* alpha,
                          int32_t limits[3]) {
    *alpha = std::sqrt(-std::log(2.0 * tolerance)) / cutoff;
    double al = *alpha;
    for (int a = 0; a < 3; ++a) {
        double width = widths[a];
        auto error = [&](long long k) {
            double t = (double)k * kPi / (width * al);
            return tolerance - 0.05 * std::sqrt(width * al) * (double)k * std::exp(-(t * t));
        };
        long long k = 10;
        if (error(k) > 0.0) {
            while (k > 0 && error(k) > 0.0) --k;
            ++k;
        } else {
            while (error(k) < 0.0) ++k;
        }
        limits[a] = (int32_t)(k | 1);
    }
}

void orc_minimum_image(double delta[3], const double box[9]) { min_image(delta, box); }

double orc_reciprocal_box(const double box[9], double recip[9]) { return reciprocal_box(box, recip); }

void orc_spline(double fraction, double weights[5], double derivatives[5]) {
    spline(fraction, weights, derivatives);
}

void orc_fft(double* data, int32_t n, int32_t inverse) {
    FftPlan plan(n);
    plan.run(reinterpret_cast<cd*>(data), inverse != 0);
}

void orc_bspline_moduli(int32_t n, double* out) {
    auto v = bspline_moduli(n);
    std::copy(v.begin(), v.end(), out);
}

double orc_pme_mesh(const double* pos, const double* amplitudes, int32_t n, const double box[9],
                    double alpha, const int32_t grid[3], int32_t dispersion, double* forces) {
    int dims[3] = {grid[0], grid[1], grid[2]};
    PmeWorkspace ws(dims);
    return ws.evaluate(pos, amplitudes, n, box, alpha, dispersion != 0, forces);
}

void orc_pme_charge_derivatives(const double* pos, const double* charges, int32_t n, const double box[9],
                                double alpha, const int32_t grid[3], double* derivatives) {
    int dims[3] = {grid[0], grid[1], grid[2]};
    PmeWorkspace ws(dims);
    ws.potential(pos, charges, n, box, alpha, false);
    double recip[9];
    reciprocal_box(box, recip);
    int nx = dims[0], ny = dims[1], nz = dims[2];
    for (int a = 0; a < n; ++a) {
        int idx[3];
        double w[3][5], dw[3][5];
        stencil(pos + 3 * a, recip, dims, idx, w, dw);
        double d = 0.0;
        for (int ix = 0; ix < 5; ++ix)
            for (int iy = 0; iy < 5; ++iy)
                for (int iz = 0; iz < 5; ++iz) {
                    size_t index =
                        ((size_t)((idx[0] + ix) % nx) * ny + (idx[1] + iy) % ny) * nz + (idx[2] + iz) % nz;
                    d += ws.grid[index].real() * w[0][ix] * w[1][iy] * w[2][iz];
                }
        derivatives[a] = d;
    }
}

// ewald.rs:49-122
double orc_ewald_reciprocal(const double* pos, const double* charges, int32_t n, const double widths[3],
                            double alpha, const int32_t limits[3], double* forces) {
    int maximum = std::max(limits[0], std::max(limits[1], limits[2]));
    std::vector<cd> phases((size_t)maximum * n * 3), xy(n), xyz(n);
    double waves[3];
    for (int a = 0; a < 3; ++a) waves[a] = 2.0 * kPi / widths[a];
    for (int i = 0; i < n; ++i)
        for (int axis = 0; axis < 3; ++axis) {
            phases[(size_t)i * 3 + axis] = cd(1.0, 0.0);
            double t = pos[3 * i + axis] * waves[axis];
            cd base(std::cos(t), std::sin(t));
            for (int k = 1; k < maximum; ++k)
                phases[((size_t)k * n + i) * 3 + axis] = phases[((size_t)(k - 1) * n + i) * 3 + axis] * base;
        }
    auto phase = [&](long long k, int i, int axis) {
        cd v = phases[((size_t)std::llabs(k) * n + i) * 3 + axis];
        return k < 0 ? std::conj(v) : v;
    };
    double coefficient = ONE_4PI_EPS0 * 4.0 * kPi / (widths[0] * widths[1] * widths[2]);
    double energy = 0.0;
    for (long long x = 0; x < limits[0]; ++x) {
        long long low_y = x == 0 ? 0 : 1 - (long long)limits[1];
        for (long long y = low_y; y < limits[1]; ++y) {
            for (int i = 0; i < n; ++i) xy[i] = phase(x, i, 0) * phase(y, i, 1);
            long long low_z = (x == 0 && y == 0) ? 1 : 1 - (long long)limits[2];
            for (long long z = low_z; z < limits[2]; ++z) {
                cd structure(0.0, 0.0);
                for (int i = 0; i < n; ++i) {
                    xyz[i] = xy[i] * phase(z, i, 2) * charges[i];
                    structure += xyz[i];
                }
                double wave[3] = {(double)x * waves[0], (double)y * waves[1], (double)z * waves[2]};
                double k2 = wave[0] * wave[0] + wave[1] * wave[1] + wave[2] * wave[2];
                double influence = coefficient * std::exp(-k2 / (4.0 * alpha * alpha)) / k2;
                energy += influence * (structure.real() * structure.real() + structure.imag() * structure.imag());
                if (forces)
                    for (int i = 0; i < n; ++i) {
                        double scale =
                            2.0 * influence * (structure.real() * xyz[i].imag() - structure.imag() * xyz[i].real());
                        for (int axis = 0; axis < 3; ++axis) forces[3 * i + axis] += scale * wave[axis];
                    }
            }
        }
    }
    return energy;
}

void orc_dispersion_reciprocal_pair(double c6, double r2, double alpha, double* energy, double* force_over_r) {
    dispersion_reciprocal_pair(c6, r2, alpha, *energy, *force_over_r);
}

double orc_dispersion_correction(const omm_nb_desc* desc, double volume) {
    Desc d(desc);
    return dispersion_coefficient(d) / volume;
}

int64_t orc_cell_candidates(const double* pos, int32_t n, double cutoff, const double* box, int32_t* out_i,
                            int32_t* out_j, int64_t capacity) {
    int64_t count = 0;
    cell_visit(pos, n, cutoff, box, [&](int i, int j) {
        if (count < capacity && out_i && out_j) {
            out_i[count] = i;
            out_j[count] = j;
        }
        ++count;
    });
    return count;
}

int64_t orc_nb_pairs(const omm_nb_desc* desc, const double* pos, const double* box, int32_t* out_i,
                     int32_t* out_j, int64_t capacity, uint64_t* hash) {
    Desc d(desc);
    int n = (int)d.particles.size();
    std::unordered_set<uint64_t, PairKeyHash> excluded;
    for (const auto& e : d.exceptions) {
        int a = std::min(e.p1, e.p2), b = std::max(e.p1, e.p2);
        excluded.insert(((uint64_t)a << 32) | (uint64_t)b);
    }
    bool use_cutoff = d.method != OMM_NB_NO_CUTOFF;
    bool use_pbc = d.periodic();
    double cutoff2 = d.cutoff * d.cutoff;
    int64_t count = 0;
    uint64_t h = 0;
    auto visit = [&](int i, int j) {
        if (excluded.count(((uint64_t)i << 32) | (uint64_t)j)) return;
        double delta[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1], pos[j * 3 + 2] - pos[i * 3 + 2]};
        if (use_pbc && box) min_image(delta, box);
        double r2 = delta[0] * delta[0];
        r2 += delta[1] * delta[1];
        r2 += delta[2] * delta[2];
        if (use_cutoff && r2 >= cutoff2) return;
        if (count < capacity && out_i && out_j) {
            out_i[count] = i;
            out_j[count] = j;
        }
        ++count;
        h += pair_key(i, j);
    };
    if (use_cutoff && n > 32)
        cell_visit(pos, n, d.cutoff, use_pbc ? box : nullptr, visit);
    else
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) visit(i, j);
    if (hash) *hash = h;
    return count;
}

int orc_nb_evaluate(const omm_nb_desc* desc, double ljpme_alpha_in, const int32_t* ljpme_grid_in,
                    const double* pos, const double* box, int include_direct, int include_reciprocal,
                    int use_cells, double* forces, double* energy_out, double parts[5]) {
    Desc d(desc);
    int n = (int)d.particles.size();
    const auto& particles = d.particles;
    const auto& exceptions = d.exceptions;
    std::vector<double> F((size_t)n * 3, 0.0);
    double energy = 0.0;
    double part[5] = {0, 0, 0, 0, 0};

    // nonbonded.rs:795-800 — exclusion set keyed by (min, max).
    std::unordered_set<uint64_t, PairKeyHash> excluded;
    for (const auto& e : exceptions) {
        int a = std::min(e.p1, e.p2), b = std::max(e.p1, e.p2);
        excluded.insert(((uint64_t)a << 32) | (uint64_t)b);
    }
    bool use_cutoff = d.method != OMM_NB_NO_CUTOFF;
    bool use_pbc = d.periodic();
    double cutoff2 = d.cutoff * d.cutoff;
    if (use_pbc) {
        if (!box) return 1;
        for (int a = 0; a < 3; ++a)
            if (!(2.0 * d.cutoff <= box[a * 3 + a] + 1e-12)) return 2;
    }
    if (use_cutoff && d.switching && !(d.switching_distance >= 0.0 && d.switching_distance < d.cutoff)) return 1;

    // nonbonded.rs:819-840 — reciprocal parameters.
    bool has_pme = false, has_ljpme = false;
    double alpha = 0.0, lj_alpha = 0.0;
    int32_t grid[3] = {0, 0, 0}, lj_grid[3] = {0, 0, 0};
    if (d.method == OMM_NB_EWALD) {
        if (d.pme_alpha > 0.0) {
            alpha = d.pme_alpha;
            for (int a = 0; a < 3; ++a) grid[a] = d.pme_grid[a];
        } else {
            if (box[3] != 0.0 || box[6] != 0.0 || box[7] != 0.0) return 1;
            double widths[3] = {box[0], box[4], box[8]};
            orc_ewald_parameters(widths, d.cutoff, d.tolerance, &alpha, grid);
        }
        has_pme = true;
    } else if (d.method == OMM_NB_PME || d.method == OMM_NB_LJPME) {
        if (d.pme_alpha > 0.0) {
            alpha = d.pme_alpha;
            for (int a = 0; a < 3; ++a) grid[a] = d.pme_grid[a];
        } else {
            pme_params(d.tolerance, d.cutoff, box, 2.0, &alpha, grid);
        }
        has_pme = true;
    }
    if (d.method == OMM_NB_LJPME) {
        if (ljpme_alpha_in > 0.0) {
            lj_alpha = ljpme_alpha_in;
            for (int a = 0; a < 3; ++a) lj_grid[a] = ljpme_grid_in[a];
        } else {
            pme_params(d.tolerance, d.cutoff, box, 1.0, &lj_alpha, lj_grid);
        }
        has_ljpme = true;
    }

    // nonbonded.rs:842-924
    auto evaluate_pair = [&](int i, int j) {
        if (excluded.count(((uint64_t)i << 32) | (uint64_t)j)) return;
        double delta[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1], pos[j * 3 + 2] - pos[i * 3 + 2]};
        if (use_pbc && box) min_image(delta, box);
        double r2 = delta[0] * delta[0];
        r2 += delta[1] * delta[1];
        r2 += delta[2] * delta[2];
        if (use_cutoff && r2 >= cutoff2) return;
        double r = std::sqrt(r2);
        const Particle& pi = particles[i];
        const Particle& pj = particles[j];
        double lj_e, lj_f;
        lj_pair(0.5 * (pi.sigma + pj.sigma), std::sqrt(pi.epsilon * pj.epsilon), r2, lj_e, lj_f);
        if (has_ljpme) {
            double c6 = 4.0 * std::sqrt(pi.epsilon * pj.epsilon) * std::pow(pi.sigma * pj.sigma, 3);
            double corr, deriv;
            dispersion_reciprocal_pair(c6, r2, lj_alpha, corr, deriv);
            double sh_e, sh_f, dc_e, dc_f;
            lj_pair(0.5 * (pi.sigma + pj.sigma), std::sqrt(pi.epsilon * pj.epsilon), cutoff2, sh_e, sh_f);
            dispersion_reciprocal_pair(c6, cutoff2, lj_alpha, dc_e, dc_f);
            double shift = sh_e + dc_e;
            lj_e += corr - shift;
            lj_f += deriv;
        }
        if (use_cutoff && d.switching && !has_ljpme && r > d.switching_distance) {
            double t = (r - d.switching_distance) / (d.cutoff - d.switching_distance);
            double sw = 1.0 + std::pow(t, 3) * (-10.0 + t * (15.0 - 6.0 * t));
            double derivative = -30.0 * t * t * std::pow(1.0 - t, 2) / (d.cutoff - d.switching_distance);
            lj_f = sw * lj_f + lj_e * derivative / r;
            lj_e *= sw;
        }
        double charge = pi.charge * pj.charge;
        double el_e, el_f;
        if (has_pme)
            screened_coulomb(charge, r, r2, alpha, el_e, el_f);
        else if (use_cutoff)
            reaction_field(charge, r, r2, d.cutoff, d.rf_dielectric, el_e, el_f);
        else
            coulomb_pair(charge, r, r2, el_e, el_f);
        double e = lj_e + el_e;
        double f_over_r = lj_f + el_f;
        energy += e;
        part[0] += e;
        for (int dd = 0; dd < 3; ++dd) {
            double f = f_over_r * delta[dd];
            F[(size_t)i * 3 + dd] += f;
            F[(size_t)j * 3 + dd] -= f;
        }
    };
    if (include_direct) {
        if (use_cutoff && use_cells && n > 32)
            cell_visit(pos, n, d.cutoff, use_pbc ? box : nullptr, evaluate_pair);
        else
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) evaluate_pair(i, j);
    }

    // nonbonded.rs:942-990
    if (include_reciprocal) {
        if (has_pme) {
            double recip[9];
            double volume = reciprocal_box(box, recip);
            double e_mesh;
            std::vector<double> charges(n);
            for (int i = 0; i < n; ++i) charges[i] = particles[i].charge;
            if (d.method == OMM_NB_EWALD) {
                double widths[3] = {box[0], box[4], box[8]};
                e_mesh = orc_ewald_reciprocal(pos, charges.data(), n, widths, alpha, grid, F.data());
            } else {
                int dims[3] = {grid[0], grid[1], grid[2]};
                PmeWorkspace ws(dims);
                e_mesh = ws.evaluate(pos, charges.data(), n, box, alpha, false, F.data());
            }
            energy += e_mesh;
            part[2] += e_mesh;
            double charge_squared = 0.0, total_charge = 0.0;
            for (const auto& p : particles) charge_squared += p.charge * p.charge;
            for (const auto& p : particles) total_charge += p.charge;
            double self = ONE_4PI_EPS0 * (alpha / std::sqrt(kPi) * charge_squared +
                                          kPi * total_charge * total_charge / (2.0 * alpha * alpha * volume));
            energy -= self;
            part[3] -= self;
        }
        if (has_ljpme) {
            std::vector<double> amp(n);
            for (int i = 0; i < n; ++i)
                amp[i] = 2.0 * std::sqrt(particles[i].epsilon) * std::pow(particles[i].sigma, 3);
            int dims[3] = {lj_grid[0], lj_grid[1], lj_grid[2]};
            PmeWorkspace ws(dims);
            double e_mesh = ws.evaluate(pos, amp.data(), n, box, lj_alpha, true, F.data());
            energy += e_mesh;
            part[2] += e_mesh;
            double s = 0.0;
            for (const auto& p : particles) s += p.epsilon * std::pow(p.sigma, 6);
            double self = s * std::pow(lj_alpha, 6) / 3.0;
            energy += self;
            part[3] += self;
        }
    }
    if (include_direct && use_pbc && d.dispersion_correction && !has_ljpme) {
        double recip[9];
        double volume = reciprocal_box(box, recip);
        double e = dispersion_coefficient(d) / volume;
        energy += e;
        part[4] += e;
    }

    // nonbonded.rs:995-1036
    if (include_direct) {
        for (const auto& exc : exceptions) {
            int i = exc.p1, j = exc.p2;
            double delta[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                               pos[j * 3 + 2] - pos[i * 3 + 2]};
            if (use_pbc && d.exceptions_pbc && box) min_image(delta, box);
            double r2 = delta[0] * delta[0];
            r2 += delta[1] * delta[1];
            r2 += delta[2] * delta[2];
            double r = std::sqrt(r2);
            double e_lj, f_lj, e_c, f_c;
            lj_pair(exc.sigma, exc.epsilon, r2, e_lj, f_lj);
            coulomb_pair(exc.charge_prod, r, r2, e_c, f_c);
            double e = e_lj + e_c, f_over_r = f_lj + f_c;
            if (has_pme) {
                double charge = particles[i].charge * particles[j].charge;
                double corr, deriv;
                reciprocal_pair(charge, r, r2, alpha, corr, deriv);
                e -= corr;
                f_over_r -= deriv;
            }
            if (has_ljpme) {
                const Particle& pi = particles[i];
                const Particle& pj = particles[j];
                double c6 = 4.0 * std::sqrt(pi.epsilon * pj.epsilon) * std::pow(pi.sigma * pj.sigma, 3);
                double corr, deriv;
                dispersion_reciprocal_pair(c6, r2, lj_alpha, corr, deriv);
                e += corr;
                f_over_r += deriv;
            }
            energy += e;
            part[1] += e;
            for (int dd = 0; dd < 3; ++dd) {
                double f = f_over_r * delta[dd];
                F[(size_t)i * 3 + dd] += f;
                F[(size_t)j * 3 + dd] -= f;
            }
        }
    }
    if (forces) std::copy(F.begin(), F.end(), forces);
    if (energy_out) *energy_out = energy;
    if (parts)
        for (int k = 0; k < 5; ++k) parts[k] = part[k];
    return 0;
}

// harmonic_bond.rs:102-144
double orc_harmonic_bonds(int32_t nb, const int32_t* p1, const int32_t* p2, const double* length,
                          const double* k, const double* pos, const double* box, int32_t use_pbc,
                          double* forces) {
    double energy = 0.0;
    for (int b = 0; b < nb; ++b) {
        int i = p1[b], j = p2[b];
        double delta[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1], pos[j * 3 + 2] - pos[i * 3 + 2]};
        if (use_pbc && box) min_image(delta, box);
        double r = std::sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
        double dr = r - length[b];
        energy += 0.5 * k[b] * dr * dr;
        double f_mag = k[b] * dr / r;
        for (int dd = 0; dd < 3; ++dd) {
            double f = f_mag * delta[dd];
            forces[i * 3 + dd] += f;
            forces[j * 3 + dd] -= f;
        }
    }
    return energy;
}

// harmonic_angle.rs:102-174
double orc_harmonic_angles(int32_t na, const int32_t* p1, const int32_t* p2, const int32_t* p3,
                           const double* theta0, const double* k, const double* pos, const double* box,
                           int32_t use_pbc, double* forces) {
    double energy = 0.0;
    auto dot3 = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    for (int a = 0; a < na; ++a) {
        int i = p1[a], j = p2[a], kk = p3[a];
        double rji[3] = {pos[i * 3] - pos[j * 3], pos[i * 3 + 1] - pos[j * 3 + 1], pos[i * 3 + 2] - pos[j * 3 + 2]};
        double rjk[3] = {pos[kk * 3] - pos[j * 3], pos[kk * 3 + 1] - pos[j * 3 + 1],
                         pos[kk * 3 + 2] - pos[j * 3 + 2]};
        if (use_pbc && box) {
            min_image(rji, box);
            min_image(rjk, box);
        }
        double rji_len = std::sqrt(dot3(rji, rji));
        double rjk_len = std::sqrt(dot3(rjk, rjk));
        double cos_theta = std::min(std::max(dot3(rji, rjk) / (rji_len * rjk_len), -1.0), 1.0);
        double theta = std::acos(cos_theta);
        double d_theta = theta - theta0[a];
        energy += 0.5 * k[a] * d_theta * d_theta;
        double sin_theta = std::max(std::sqrt(1.0 - cos_theta * cos_theta), 1e-8);
        double de_dtheta = k[a] * d_theta;
        for (int dd = 0; dd < 3; ++dd) {
            double fi = de_dtheta * (cos_theta * rji[dd] / rji_len - rjk[dd] / rjk_len) / (rji_len * sin_theta);
            double fk = de_dtheta * (cos_theta * rjk[dd] / rjk_len - rji[dd] / rji_len) / (rjk_len * sin_theta);
            forces[i * 3 + dd] -= fi;
            forces[kk * 3 + dd] -= fk;
            forces[j * 3 + dd] += fi + fk;
        }
    }
    return energy;
}

double orc_boltzmann(void) { return BOLTZMANN; }

// langevin.rs:103-142
void orc_lfmiddle_step(int32_t n, double* pos, double* vel, const double* forces, const double* masses,
                       double dt, double temperature, double friction, const double* normals) {
    double half_dt = 0.5 * dt;
    double decay = std::exp(-friction * dt);
    double variance = BOLTZMANN * temperature * -std::expm1(-2.0 * friction * dt);
    size_t r = 0;
    for (int atom = 0; atom < n; ++atom) {
        double mass = masses[atom];
        if (mass == 0.0) continue;
        double inverse_mass = 1.0 / mass;
        double noise_scale = std::sqrt(variance * inverse_mass);
        for (int axis = 0; axis < 3; ++axis) {
            size_t i = (size_t)3 * atom + axis;
            vel[i] += forces[i] * inverse_mass * dt;
            double midpoint = pos[i] + half_dt * vel[i];
            vel[i] = decay * vel[i] + noise_scale * (normals ? normals[r] : 0.0);
            ++r;
            pos[i] = midpoint + half_dt * vel[i];
        }
    }
}

// constraints.rs:35-149 — sparse SHAKE (positions) / RATTLE (velocities)
// projection: Gauss-Seidel sweeps over the constraints in order, directions
// from the reference positions, until every constraint is within tolerance
// (at most 1000 sweeps).  Constraints between two massless particles are
// dropped; returns 0, or 1 on a reference error (invalid input, singular or
// nonfinite projection, no convergence).
int32_t orc_constraints_solve(int32_t n, int32_t nc, const int32_t* c1, const int32_t* c2, const double* dist,
                              const double* masses, const double* reference, double* values, int32_t velocities,
                              double tolerance) {
    if (!(std::isfinite(tolerance) && tolerance > 0.0 && tolerance < 1.0)) return 1;
    struct Prepared {
        int i, j;
        double wi, wj, d2;
    };
    std::vector<Prepared> cs;
    for (int k = 0; k < nc; ++k) {
        const int i = c1[k], j = c2[k];
        if (i < 0 || j < 0 || i >= n || j >= n || i == j) return 1;
        if (!(std::isfinite(dist[k]) && dist[k] > 0.0)) return 1;
        if ((masses[i] == 0.0) != (masses[j] == 0.0)) return 1;
        if (masses[i] != 0.0) cs.push_back({i, j, 1.0 / masses[i], 1.0 / masses[j], dist[k] * dist[k]});
    }
    std::vector<std::array<double, 3>> dir(cs.size());
    for (size_t k = 0; k < cs.size(); ++k)
        for (int a = 0; a < 3; ++a) dir[k][a] = reference[3 * cs[k].i + a] - reference[3 * cs[k].j + a];
    for (int sweep = 0; sweep < 1000; ++sweep) {
        bool converged = true;
        for (size_t k = 0; k < cs.size(); ++k) {
            const Prepared& c = cs[k];
            double delta[3], dot = 0.0;
            for (int a = 0; a < 3; ++a) {
                delta[a] = values[3 * c.i + a] - values[3 * c.j + a];
                dot += delta[a] * dir[k][a];
            }
            double mult;
            if (velocities) {
                double norm = 0.0;
                for (int a = 0; a < 3; ++a) norm += dir[k][a] * dir[k][a];
                if (norm == 0.0 || !std::isfinite(norm)) return 1;
                const double correction = -dot / ((c.wi + c.wj) * norm);
                if (std::fabs(correction) <= tolerance) continue;
                mult = correction;
            } else {
                double norm = 0.0;
                for (int a = 0; a < 3; ++a) norm += delta[a] * delta[a];
                const double ratio = std::sqrt(norm / c.d2);
                if (std::isfinite(ratio) && std::fabs(ratio - 1.0) <= tolerance) continue;
                const double denominator = 2.0 * (c.wi + c.wj) * dot;
                if (std::fabs(denominator) < DBL_MIN || !std::isfinite(denominator)) return 1;
                mult = (c.d2 - norm) / denominator;
            }
            if (!std::isfinite(mult)) return 1;
            converged = false;
            for (int a = 0; a < 3; ++a) {
                const double correction = mult * dir[k][a];
                values[3 * c.i + a] += c.wi * correction;
                values[3 * c.j + a] -= c.wj * correction;
            }
        }
        if (converged) return 0;
    }
    return 1;
}

// langevin.rs:215-277 (host path) — constrained LFMiddle: kick, RATTLE the
// velocities against the current positions, drift / friction + noise / drift,
// SHAKE the positions against the previous ones, and add the position
// correction / dt to the velocities.  normals as in orc_lfmiddle_step.
int32_t orc_lfmiddle_step_constrained(int32_t n, double* pos, double* vel, const double* forces,
                                      const double* masses, double dt, double temperature, double friction,
                                      const double* normals, int32_t nc, const int32_t* c1, const int32_t* c2,
                                      const double* dist, double tolerance) {
    const std::vector<double> previous(pos, pos + 3 * (size_t)n);
    for (int atom = 0; atom < n; ++atom) {
        if (masses[atom] == 0.0) continue;
        const double inverse_mass = 1.0 / masses[atom];
        for (int axis = 0; axis < 3; ++axis) vel[3 * atom + axis] += forces[3 * atom + axis] * inverse_mass * dt;
    }
    if (orc_constraints_solve(n, nc, c1, c2, dist, masses, pos, vel, 1, tolerance)) return 1;
    const double half_dt = 0.5 * dt;
    const double decay = std::exp(-friction * dt);
    const double variance = BOLTZMANN * temperature * -std::expm1(-2.0 * friction * dt);
    size_t r = 0;
    for (int atom = 0; atom < n; ++atom) {
        if (masses[atom] == 0.0) continue;
        const double noise_scale = std::sqrt(variance / masses[atom]);
        for (int axis = 0; axis < 3; ++axis) {
            const size_t i = (size_t)3 * atom + axis;
            const double midpoint = pos[i] + half_dt * vel[i];
            vel[i] = decay * vel[i] + noise_scale * (normals ? normals[r] : 0.0);
            ++r;
            pos[i] = midpoint + half_dt * vel[i];
        }
    }
    const std::vector<double> proposed(pos, pos + 3 * (size_t)n);
    if (orc_constraints_solve(n, nc, c1, c2, dist, masses, previous.data(), pos, 0, tolerance)) return 1;
    for (int atom = 0; atom < n; ++atom) {
        if (masses[atom] == 0.0) continue;
        for (int axis = 0; axis < 3; ++axis) {
            const size_t i = (size_t)3 * atom + axis;
            vel[i] += (pos[i] - proposed[i]) / dt;
        }
    }
    return 0;
}

// system.rs:100-141 — union-find over bonded pairs.
int32_t orc_molecules(int32_t n, int32_t nbonds, const int32_t* p1, const int32_t* p2, int32_t* molecule_of) {
    std::vector<int> parents(n);
    for (int i = 0; i < n; ++i) parents[i] = i;
    auto root = [&](int p) {
        while (parents[p] != p) {
            parents[p] = parents[parents[p]];
            p = parents[p];
        }
        return p;
    };
    for (int b = 0; b < nbonds; ++b) {
        int a = root(p1[b]), c = root(p2[b]);
        parents[std::max(a, c)] = std::min(a, c);
    }
    std::vector<int> id(n, -1);
    int count = 0;
    for (int i = 0; i < n; ++i) {
        int r0 = root(i);
        if (id[r0] < 0) id[r0] = count++;
        molecule_of[i] = id[r0];
    }
    return count;
}

// context.rs:976-1029 for transform = diag(s): rigid moves each molecule by
// (s-1) * centroid; otherwise every coordinate is scaled.
void orc_scale_positions(int32_t n, double* pos, double scale, int32_t rigid, const int32_t* molecule_of,
                         int32_t nmol) {
    if (!rigid) {
        for (int i = 0; i < 3 * n; ++i) pos[i] = scale * pos[i];
        return;
    }
    double t = scale - 1.0;
    std::vector<double> center((size_t)nmol * 3, 0.0);
    std::vector<int> count(nmol, 0);
    // Sum members in increasing atom order, as the molecule lists are.
    for (int i = 0; i < n; ++i) {
        int m = molecule_of[i];
        for (int a = 0; a < 3; ++a) center[(size_t)m * 3 + a] += pos[3 * i + a];
        ++count[m];
    }
    for (int m = 0; m < nmol; ++m) {
        double inverse_count = 1.0 / (double)count[m];
        for (int a = 0; a < 3; ++a) center[(size_t)m * 3 + a] *= inverse_count;
    }
    for (int i = 0; i < n; ++i) {
        int m = molecule_of[i];
        for (int a = 0; a < 3; ++a) pos[3 * i + a] += t * center[(size_t)m * 3 + a];
    }
}

// barostat.rs:160-163
// flexible_barostat.rs:15-41: flip vectors with a negative diagonal, then
// reduce c by b (y), c by a (x) and b by a (x).  0 = a valid reduced box.
int32_t orc_reduced_box(double* v) {
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(v[k])) return 1;
    for (int row = 0; row < 3; ++row) {
        if (v[3 * row + row] == 0.0) return 1;
        if (v[3 * row + row] < 0.0)
            for (int c = 0; c < 3; ++c) v[3 * row + c] = -v[3 * row + c];
    }
    const int steps[3][3] = {{2, 1, 1}, {2, 0, 0}, {1, 0, 0}};
    for (const auto& st : steps) {
        const double shift = std::round(v[3 * st[0] + st[2]] / v[3 * st[1] + st[2]]);
        for (int c = 0; c < 3; ++c) v[3 * st[0] + c] -= shift * v[3 * st[1] + c];
    }
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(v[k])) return 1;
    return 0;
}

// context.rs:976-1029 with a diagonal transform (the anisotropic barostat's
// proposal, anisotropic_barostat.rs:148-157): rigid molecules move by
// (s_a - 1) times their unweighted centroid along each axis a.
void orc_scale_positions_axes(int32_t n, double* pos, const double scale[3], const int32_t* molecule_of,
                              int32_t nmol) {
    std::vector<double> center((size_t)nmol * 3, 0.0);
    std::vector<int> count(nmol, 0);
    for (int i = 0; i < n; ++i) {
        int m = molecule_of[i];
        for (int a = 0; a < 3; ++a) center[(size_t)m * 3 + a] += pos[3 * i + a];
        ++count[m];
    }
    for (int m = 0; m < nmol; ++m) {
        double inverse_count = 1.0 / (double)count[m];
        for (int a = 0; a < 3; ++a) center[(size_t)m * 3 + a] *= inverse_count;
    }
    for (int i = 0; i < n; ++i) {
        int m = molecule_of[i];
        for (int a = 0; a < 3; ++a) pos[3 * i + a] += (scale[a] - 1.0) * center[(size_t)m * 3 + a];
    }
}

double orc_barostat_work(double energy_change, double pressure_bar, double old_volume, double new_volume,
                         int32_t scaled_particles, double temperature) {
    double kt = BOLTZMANN * temperature;
    return energy_change + pressure_bar * BAR_NM3_TO_KJ_PER_MOL * (new_volume - old_volume) -
           (double)scaled_particles * kt * std::log(new_volume / old_volume);
}

}  // extern "C"
