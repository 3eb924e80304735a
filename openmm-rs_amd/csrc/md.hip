// Device-resident MD step for the flexible-water workloads: NonbondedForce
// (PME) + HarmonicBondForce + HarmonicAngleForce + LangevinMiddle, captured
// into a HIP graph (Context::step, context.rs:795-928).
#include <cmath>
#include <random>

#include "md.h"

namespace ommhip {

namespace {

// BOLTZMANN in kJ/mol/K (langevin.rs:12).
const double BOLTZMANN = 1.380649e-23 * 6.02214076e23 / 1000.0;
constexpr int GRAPH_STEPS = 10;

// HarmonicBondForce without PBC (harmonic_bond.rs:113-140).
template <class A>
__global__ void k_bonds(int nb, const int* __restrict__ p1, const int* __restrict__ p2,
                        const double* __restrict__ length, const double* __restrict__ kb,
                        const double* __restrict__ pos, A* force, A* energy, unsigned* status) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (b < nb) {
        int i = p1[b], j = p2[b];
        double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double dr = r - length[b];
        e = 0.5 * kb[b] * dr * dr;
        double f = kb[b] * dr / r;
        for (int a = 0; a < 3; ++a) {
            acc_add(force + 3 * i + a, f * d[a], FORCE_SCALE, status);
            acc_add(force + 3 * j + a, -f * d[a], FORCE_SCALE, status);
        }
    }
    double et = wave_sum(e);
    if ((threadIdx.x & 63) == 0) acc_add(energy + E_BOND, et, ENERGY_SCALE, status);
}

// HarmonicAngleForce without PBC (harmonic_angle.rs:118-167).
template <class A>
__global__ void k_angles(int na, const int* __restrict__ p1, const int* __restrict__ p2, const int* __restrict__ p3,
                         const double* __restrict__ theta0, const double* __restrict__ ka,
                         const double* __restrict__ pos, A* force, A* energy, unsigned* status) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (t < na) {
        int i = p1[t], j = p2[t], k = p3[t];
        double rji[3], rjk[3];
        for (int a = 0; a < 3; ++a) {
            rji[a] = pos[3 * i + a] - pos[3 * j + a];
            rjk[a] = pos[3 * k + a] - pos[3 * j + a];
        }
        double lji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
        double ljk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
        double c = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (lji * ljk);
        c = fmin(fmax(c, -1.0), 1.0);
        double theta = acos(c);
        double dth = theta - theta0[t];
        e = 0.5 * ka[t] * dth * dth;
        double s = fmax(sqrt(1.0 - c * c), 1e-8);
        double de = ka[t] * dth;
        for (int a = 0; a < 3; ++a) {
            double fi = de * (c * rji[a] / lji - rjk[a] / ljk) / (lji * s);
            double fk = de * (c * rjk[a] / ljk - rji[a] / lji) / (ljk * s);
            acc_add(force + 3 * i + a, -fi, FORCE_SCALE, status);
            acc_add(force + 3 * k + a, -fk, FORCE_SCALE, status);
            acc_add(force + 3 * j + a, fi + fk, FORCE_SCALE, status);
        }
    }
    double et = wave_sum(e);
    if ((threadIdx.x & 63) == 0) acc_add(energy + E_ANGLE, et, ENERGY_SCALE, status);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

__global__ void k_step_begin(unsigned long long* step) { step[0] += 1ull; }

inline __device__ double force_value(long long f) { return (double)f * (1.0 / FORCE_SCALE); }
inline __device__ double force_value(double f) { return f; }

// LangevinMiddle update (langevin.rs:103-142) with counter-based normals.
template <class A>
__global__ void k_lfmiddle(int n, double* __restrict__ pos, double* __restrict__ vel, const A* __restrict__ force,
                           const double* __restrict__ inv_mass, double dt, double decay, double variance,
                           uint64_t seed, const unsigned long long* __restrict__ step) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double im = inv_mass[i];
    if (im == 0.0) return;  // massless particles do not move
    uint32_t c[4] = {(uint32_t)i, (uint32_t)step[0], (uint32_t)(step[0] >> 32), 0x4c464d31u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = ((double)c[0] + 1.0) * (1.0 / 4294967296.0);  // (0, 1]
    double u2 = (double)c[1] * (1.0 / 4294967296.0);
    double u3 = ((double)c[2] + 1.0) * (1.0 / 4294967296.0);
    double u4 = (double)c[3] * (1.0 / 4294967296.0);
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    double xi[3] = {r1 * cos(2.0 * PI * u2), r1 * sin(2.0 * PI * u2), r2 * cos(2.0 * PI * u4)};
    double noise = sqrt(variance * im);
    double half = 0.5 * dt;
    for (int a = 0; a < 3; ++a) {
        int k = 3 * i + a;
        double v = vel[k] + force_value(force[k]) * im * dt;
        double mid = pos[k] + half * v;
        v = decay * v + noise * xi[a];
        vel[k] = v;
        pos[k] = mid + half * v;
    }
}

}  // namespace

MdContext::MdContext(const omm_md_desc* d, int precision) {
    if (!d || !d->nonbonded) throw Error(OMM_ERR_INVALID, "null MD descriptor");
    n_ = d->num_particles;
    if (n_ <= 0 || !d->masses) throw Error(OMM_ERR_INVALID, "MD descriptor needs particles and masses");
    if (d->nonbonded->num_particles != n_) throw Error(OMM_ERR_INVALID, "NonbondedForce particle count mismatch");
    if (!(std::isfinite(d->temperature) && d->temperature >= 0.0))
        throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
    if (!(std::isfinite(d->friction) && d->friction >= 0.0))
        throw Error(OMM_ERR_INVALID, "friction must be nonnegative and finite");
    if (!(std::isfinite(d->step_size) && d->step_size > 0.0))
        throw Error(OMM_ERR_INVALID, "step size must be positive and finite");
    mixed_ = precision == OMM_PRECISION_MIXED;
    nb_.reset(new NbOperator(d->nonbonded, precision));
    nb_->initialize(n_, d->box);
    s_ = nb_->stream();
    masses_.assign(d->masses, d->masses + n_);
    std::vector<double> inv(n_);
    for (int i = 0; i < n_; ++i) {
        if (!(std::isfinite(masses_[i]) && masses_[i] >= 0.0)) throw Error(OMM_ERR_INVALID, "masses must be nonnegative");
        inv[i] = masses_[i] == 0.0 ? 0.0 : 1.0 / masses_[i];
    }
    inv_mass_.upload(inv, s_);
    nbonds_ = d->num_bonds;
    nangles_ = d->num_angles;
    auto check_idx = [&](const int32_t* p, int cnt) {
        for (int i = 0; i < cnt; ++i)
            if (p[i] < 0 || p[i] >= n_) throw Error(OMM_ERR_INVALID, "bonded particle index out of range");
    };
    if (nbonds_ > 0) {
        check_idx(d->bond_particle1, nbonds_);
        check_idx(d->bond_particle2, nbonds_);
        b1_.upload(d->bond_particle1, nbonds_, s_);
        b2_.upload(d->bond_particle2, nbonds_, s_);
        blen_.upload(d->bond_length, nbonds_, s_);
        bk_.upload(d->bond_k, nbonds_, s_);
    }
    if (nangles_ > 0) {
        check_idx(d->angle_particle1, nangles_);
        check_idx(d->angle_particle2, nangles_);
        check_idx(d->angle_particle3, nangles_);
        a1_.upload(d->angle_particle1, nangles_, s_);
        a2_.upload(d->angle_particle2, nangles_, s_);
        a3_.upload(d->angle_particle3, nangles_, s_);
        ath_.upload(d->angle_theta0, nangles_, s_);
        ak_.upload(d->angle_k, nangles_, s_);
    }
    box_h_ = make_box(d->box);
    box_.upload(&box_h_, 1, s_);
    dt_ = d->step_size;
    temperature_ = d->temperature;
    friction_ = d->friction;
    seed_ = d->seed;
    pos_.alloc(3 * (size_t)n_);
    vel_.alloc(3 * (size_t)n_);
    fill_async<double>(vel_.p, 3 * (size_t)n_, 0.0, s_);
    if (mixed_) {
        force_fx_.alloc(3 * (size_t)n_);
        energy_fx_.alloc(E_SLOTS);
    } else {
        force_d_.alloc(3 * (size_t)n_);
        energy_d_.alloc(E_SLOTS);
    }
    out_.alloc(3 * (size_t)n_);
    status_.alloc(1);
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    step_dev_.alloc(1);
    fill_async<unsigned long long>(step_dev_.p, 1, 0ull, s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

MdContext::~MdContext() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
}

void MdContext::set_positions(const double* pos) {
    for (int i = 0; i < 3 * n_; ++i)
        if (!std::isfinite(pos[i])) throw Error(OMM_ERR_NONFINITE, "particle positions must be finite");
    pos_.upload(pos, 3 * (size_t)n_, s_);
    if (nb_->calibrate(pos_.p, box_.p, status_.p, s_)) drop_graph();
    OMM_HIP(hipStreamSynchronize(s_));
}

void MdContext::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

void MdContext::set_velocities(const double* vel) {
    vel_.upload(vel, 3 * (size_t)n_, s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

void MdContext::enqueue_forces(hipStream_t s, bool energy) {
    if (mixed_) {
        fill_async<long long>(force_fx_.p, 3 * (size_t)n_, 0ll, s);
        fill_async<long long>(energy_fx_.p, E_SLOTS, 0ll, s);
    } else {
        fill_async<double>(force_d_.p, 3 * (size_t)n_, 0.0, s);
        fill_async<double>(energy_d_.p, E_SLOTS, 0.0, s);
    }
    const NbSettings& c = nb_->settings();
    nb_->enqueue(pos_.p, box_.p, c.include_direct, true, energy, force_fx_.p, force_d_.p, energy_fx_.p, energy_d_.p,
                 status_.p, s);
    if (nbonds_ > 0) {
        if (mixed_)
            hipLaunchKernelGGL(k_bonds<long long>, dim3((nbonds_ + 255) / 256), dim3(256), 0, s, nbonds_, b1_.p, b2_.p,
                               blen_.p, bk_.p, pos_.p, force_fx_.p, energy_fx_.p, status_.p);
        else
            hipLaunchKernelGGL(k_bonds<double>, dim3((nbonds_ + 255) / 256), dim3(256), 0, s, nbonds_, b1_.p, b2_.p,
                               blen_.p, bk_.p, pos_.p, force_d_.p, energy_d_.p, status_.p);
    }
    if (nangles_ > 0) {
        if (mixed_)
            hipLaunchKernelGGL(k_angles<long long>, dim3((nangles_ + 255) / 256), dim3(256), 0, s, nangles_, a1_.p,
                               a2_.p, a3_.p, ath_.p, ak_.p, pos_.p, force_fx_.p, energy_fx_.p, status_.p);
        else
            hipLaunchKernelGGL(k_angles<double>, dim3((nangles_ + 255) / 256), dim3(256), 0, s, nangles_, a1_.p,
                               a2_.p, a3_.p, ath_.p, ak_.p, pos_.p, force_d_.p, energy_d_.p, status_.p);
    }
    OMM_HIP(hipGetLastError());
}

void MdContext::enqueue_step(hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, s, step_dev_.p);
    enqueue_forces(s, false);
    double decay = std::exp(-friction_ * dt_);
    double variance = BOLTZMANN * temperature_ * -std::expm1(-2.0 * friction_ * dt_);
    if (mixed_)
        hipLaunchKernelGGL(k_lfmiddle<long long>, dim3((n_ + 255) / 256), dim3(256), 0, s, n_, pos_.p, vel_.p,
                           force_fx_.p, inv_mass_.p, dt_, decay, variance, seed_, step_dev_.p);
    else
        hipLaunchKernelGGL(k_lfmiddle<double>, dim3((n_ + 255) / 256), dim3(256), 0, s, n_, pos_.p, vel_.p,
                           force_d_.p, inv_mass_.p, dt_, decay, variance, seed_, step_dev_.p);
    OMM_HIP(hipGetLastError());
}

void MdContext::check_status() {
    unsigned st = 0;
    status_.download(&st, 1, s_);
    OMM_HIP(hipStreamSynchronize(s_));
    if (st & ST_LIST_OVERFLOW) {
        fill_async<unsigned>(status_.p, 1, 0u, s_);
        throw Error(OMM_ERR_OVERFLOW, "neighbour list capacity exceeded during the step");
    }
    if (st & ST_NONFINITE) throw Error(OMM_ERR_NONFINITE, "non-finite force or energy on the device");
    if (st & ST_OVERFLOW) throw Error(OMM_ERR_OVERFLOW, "fixed-point force/energy accumulator overflow");
}

void MdContext::step(int nsteps) {
    if (nsteps < 0) throw Error(OMM_ERR_INVALID, "step count must be nonnegative");
    if (!mixed_) {
        // Plain FP64 form: one launch sequence per step, no graph.
        for (int k = 0; k < nsteps; ++k) enqueue_step(s_);
        step_count_ += nsteps;
        check_status();
        return;
    }
    if (!graph_exec_ && nsteps >= GRAPH_STEPS) {
        OMM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < GRAPH_STEPS; ++k) enqueue_step(s_);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        OMM_HIP(hipStreamEndCapture(s_, &graph_));
        OMM_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
    }
    int done = 0;
    while (graph_exec_ && nsteps - done >= GRAPH_STEPS) {
        OMM_HIP(hipGraphLaunch(graph_exec_, s_));
        done += GRAPH_STEPS;
    }
    for (; done < nsteps; ++done) enqueue_step(s_);
    step_count_ += nsteps;
    check_status();
}

double MdContext::evaluate(double* forces, bool energy) {
    enqueue_forces(s_, energy);
    double e[E_SLOTS];
    if (mixed_) {
        long long efx[E_SLOTS];
        energy_fx_.download(efx, E_SLOTS, s_);
        if (forces) {
            launch_fixed_to_double(force_fx_.p, out_.p, 3 * (size_t)n_, s_);
            out_.download(forces, 3 * (size_t)n_, s_);
        }
        OMM_HIP(hipStreamSynchronize(s_));
        for (int k = 0; k < E_SLOTS; ++k) e[k] = (double)efx[k] * (1.0 / ENERGY_SCALE);
    } else {
        energy_d_.download(e, E_SLOTS, s_);
        if (forces) force_d_.download(forces, 3 * (size_t)n_, s_);
        OMM_HIP(hipStreamSynchronize(s_));
    }
    check_status();
    const NbSettings& c = nb_->settings();
    HostTerms ht = nb_->host_terms(box_h_.volume, c.include_direct, true);
    return e[E_DIRECT] + e[E_RECIP] + ht.self + ht.dispersion + e[E_EXCEPT] + e[E_BOND] + e[E_ANGLE];
}

void MdContext::set_velocities_to_temperature(double temperature, uint64_t seed) {
    // context.rs:182-223
    if (!(std::isfinite(temperature) && temperature >= 0.0))
        throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> normal(0.0, 1.0);
    std::vector<double> v(3 * (size_t)n_), f(3 * (size_t)n_);
    for (int i = 0; i < n_; ++i) {
        double scale = masses_[i] == 0.0 ? 0.0 : std::sqrt(BOLTZMANN * temperature / masses_[i]);
        for (int a = 0; a < 3; ++a) v[3 * i + a] = normal(rng) * scale;
    }
    evaluate(f.data(), false);
    double offset = 0.5 * dt_;  // LangevinMiddle velocity_time_offset (langevin.rs:204-206)
    for (int i = 0; i < n_; ++i) {
        double im = masses_[i] == 0.0 ? 0.0 : 1.0 / masses_[i];
        for (int a = 0; a < 3; ++a) v[3 * i + a] -= f[3 * i + a] * im * offset;
    }
    set_velocities(v.data());
}

void MdContext::get_state(double* pos, double* vel, double* forces, double* energy, int64_t* step) {
    if (forces || energy) {
        double e = evaluate(forces, energy != nullptr);
        if (energy) *energy = e;
    }
    if (pos) pos_.download(pos, 3 * (size_t)n_, s_);
    if (vel) vel_.download(vel, 3 * (size_t)n_, s_);
    OMM_HIP(hipStreamSynchronize(s_));
    if (step) *step = step_count_;
}

void MdContext::profile(int reps, double ms[8], int64_t* num_pairs) {
    if (reps <= 0) throw Error(OMM_ERR_INVALID, "reps must be positive");
    nb_->profile(pos_.p, box_.p, reps, force_fx_.p, force_d_.p, energy_fx_.p, energy_d_.p, status_.p, ms, num_pairs,
                 s_);
    hipEvent_t t0, t1;
    OMM_HIP(hipEventCreate(&t0));
    OMM_HIP(hipEventCreate(&t1));
    auto timed = [&](auto&& body) {
        body();
        OMM_HIP(hipEventRecord(t0, s_));
        for (int r = 0; r < reps; ++r) body();
        OMM_HIP(hipEventRecord(t1, s_));
        OMM_HIP(hipEventSynchronize(t1));
        float v = 0.f;
        OMM_HIP(hipEventElapsedTime(&v, t0, t1));
        return (double)v / reps;
    };
    ms[6] = timed([&] {
        if (nbonds_ > 0 && mixed_)
            hipLaunchKernelGGL(k_bonds<long long>, dim3((nbonds_ + 255) / 256), dim3(256), 0, s_, nbonds_, b1_.p,
                               b2_.p, blen_.p, bk_.p, pos_.p, force_fx_.p, energy_fx_.p, status_.p);
        if (nangles_ > 0 && mixed_)
            hipLaunchKernelGGL(k_angles<long long>, dim3((nangles_ + 255) / 256), dim3(256), 0, s_, nangles_, a1_.p,
                               a2_.p, a3_.p, ath_.p, ak_.p, pos_.p, force_fx_.p, energy_fx_.p, status_.p);
        if (nbonds_ > 0 && !mixed_)
            hipLaunchKernelGGL(k_bonds<double>, dim3((nbonds_ + 255) / 256), dim3(256), 0, s_, nbonds_, b1_.p, b2_.p,
                               blen_.p, bk_.p, pos_.p, force_d_.p, energy_d_.p, status_.p);
        if (nangles_ > 0 && !mixed_)
            hipLaunchKernelGGL(k_angles<double>, dim3((nangles_ + 255) / 256), dim3(256), 0, s_, nangles_, a1_.p,
                               a2_.p, a3_.p, ath_.p, ak_.p, pos_.p, force_d_.p, energy_d_.p, status_.p);
    });
    // The integrator is timed on scratch copies of the state.
    DevBuf<double> px, vx;
    px.alloc(3 * (size_t)n_);
    vx.alloc(3 * (size_t)n_);
    OMM_HIP(hipMemcpyAsync(px.p, pos_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    OMM_HIP(hipMemcpyAsync(vx.p, vel_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    double decay = std::exp(-friction_ * dt_);
    double variance = BOLTZMANN * temperature_ * -std::expm1(-2.0 * friction_ * dt_);
    fill_async<long long>(force_fx_.p, force_fx_.n, 0ll, s_);
    fill_async<double>(force_d_.p, force_d_.n, 0.0, s_);
    ms[7] = timed([&] {
        if (mixed_)
            hipLaunchKernelGGL(k_lfmiddle<long long>, dim3((n_ + 255) / 256), dim3(256), 0, s_, n_, px.p, vx.p,
                               force_fx_.p, inv_mass_.p, dt_, decay, variance, seed_, step_dev_.p);
        else
            hipLaunchKernelGGL(k_lfmiddle<double>, dim3((n_ + 255) / 256), dim3(256), 0, s_, n_, px.p, vx.p,
                               force_d_.p, inv_mass_.p, dt_, decay, variance, seed_, step_dev_.p);
    });
    OMM_HIP(hipStreamSynchronize(s_));
    OMM_HIP(hipEventDestroy(t0));
    OMM_HIP(hipEventDestroy(t1));
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

double MdContext::kinetic_energy() {
    std::vector<double> v(3 * (size_t)n_);
    vel_.download(v.data(), v.size(), s_);
    OMM_HIP(hipStreamSynchronize(s_));
    double ke = 0.0;
    for (int i = 0; i < n_; ++i)
        ke += 0.5 * masses_[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return ke;
}

}  // namespace ommhip
