// Device-resident MD step for the flexible-water workloads: NonbondedForce
// (PME) + HarmonicBondForce + HarmonicAngleForce + LangevinMiddle, captured
// into a HIP graph (Context::step, context.rs:795-928).
#include <cmath>
#include <random>

#include "md.h"
#include "terms.h"

namespace ommhip {

namespace {

// BOLTZMANN in kJ/mol/K (langevin.rs:12).
const double BOLTZMANN = 1.380649e-23 * 6.02214076e23 / 1000.0;
constexpr int GRAPH_STEPS = 10;

// Exceptions, harmonic bonds and harmonic angles in one launch: threads
// [0, ne) take exceptions, [ne, ne+nb) bonds, the rest angles.  Indices are
// global over the batch; energies go to the term's system.
template <class A>
__global__ void k_bonded_all(int ne, const int* __restrict__ e1, const int* __restrict__ e2,
                             const double* __restrict__ eqq, const double* __restrict__ esig,
                             const double* __restrict__ eeps, const double* __restrict__ charge, int use_pbc,
                             int has_pme, double alpha, int nb, const int* __restrict__ b1, const int* __restrict__ b2,
                             const double* __restrict__ blen, const double* __restrict__ bk, int na,
                             const int* __restrict__ a1, const int* __restrict__ a2, const int* __restrict__ a3,
                             const double* __restrict__ ath, const double* __restrict__ ak,
                             const double* __restrict__ pos, const int* __restrict__ atom_sys,
                             const SysInfo* __restrict__ sys, A* force, A* energy, unsigned* status) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double ee = 0.0, eb = 0.0, ea = 0.0;
    int se = -1, sb = -1, sa = -1;
    if (t < ne) {
        int i = e1[t], j = e2[t];
        se = atom_sys[i];
        ee = exception_term(i, j, eqq[t], esig[t], eeps[t], charge[i], charge[j], pos, sys[se].box, use_pbc, has_pme,
                            alpha, force, status);
    } else if (t < ne + nb) {
        int b = t - ne;
        sb = atom_sys[b1[b]];
        eb = bond_term(b1[b], b2[b], blen[b], bk[b], pos, force, status);
    } else if (t < ne + nb + na) {
        int a = t - ne - nb;
        sa = atom_sys[a1[a]];
        ea = angle_term(a1[a], a2[a], a3[a], ath[a], ak[a], pos, force, status);
    }
    wave_energy_add(energy, se, E_EXCEPT, ee, status);
    wave_energy_add(energy, sb, E_BOND, eb, status);
    wave_energy_add(energy, sa, E_ANGLE, ea, status);
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

// First kernel of every step: advance the RNG counter, clear the energies.
template <class A>
__global__ void k_step_begin(unsigned long long* step, A* energy, int nslots) {
    if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1ull;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) energy[i] = (A)0;
}

__global__ void k_copy2(const double* __restrict__ a, double* __restrict__ a_out, const double* __restrict__ b,
                        double* __restrict__ b_out, size_t n, const unsigned long long* __restrict__ step,
                        unsigned long long* step_out) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) {
        a_out[i] = a[i];
        b_out[i] = b[i];
    }
    if (i == 0) step_out[0] = step[0];
}

inline __device__ double force_value(long long f) { return (double)f * (1.0 / FORCE_SCALE); }
inline __device__ double force_value(double f) { return f; }

// LangevinMiddle update (langevin.rs:103-142) with counter-based normals:
// Philox keyed by the system's seed, counter (index within the system,
// step), so a replica draws the same numbers alone or in a batch.  The force
// accumulator is read and cleared for the next step.  check != 0: raise the
// neighbour list's rebuild flags (any, and the atom's system) when an atom
// moved more than skin/2 from its position at the last build, so the next
// step's build kernels rebuild that system.
template <class A>
__global__ void k_lfmiddle(int n, double* __restrict__ pos, double* __restrict__ vel, A* __restrict__ force,
                           const double* __restrict__ inv_mass, const int* __restrict__ atom_sys,
                           const SysInfo* __restrict__ sys, double dt, double decay, double variance,
                           const unsigned long long* __restrict__ step, const double* __restrict__ pos_build,
                           double lim2, int check, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    int si = 0;
    if (i < n) {
        double im = inv_mass[i];
        double f[3];
        for (int a = 0; a < 3; ++a) {
            f[a] = force_value(force[3 * i + a]);
            force[3 * i + a] = (A)0;
        }
        if (im != 0.0) {  // massless particles do not move
            si = atom_sys[i];
            const SysInfo& S = sys[si];
            const uint64_t seed = S.seed;
            uint32_t c[4] = {(uint32_t)(i - S.atom_off), (uint32_t)step[0], (uint32_t)(step[0] >> 32), 0x4c464d31u};
            philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            double u1 = ((double)c[0] + 1.0) * (1.0 / 4294967296.0);  // (0, 1]
            double u2 = (double)c[1] * (1.0 / 4294967296.0);
            double u3 = ((double)c[2] + 1.0) * (1.0 / 4294967296.0);
            double u4 = (double)c[3] * (1.0 / 4294967296.0);
            double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
            double xi[3] = {r1 * cos(2.0 * PI * u2), r1 * sin(2.0 * PI * u2), r2 * cos(2.0 * PI * u4)};
            double noise = sqrt(variance * im);
            double half = 0.5 * dt;
            double d2 = 0.0;
            for (int a = 0; a < 3; ++a) {
                int k = 3 * i + a;
                double v = vel[k] + f[a] * im * dt;
                double mid = pos[k] + half * v;
                v = decay * v + noise * xi[a];
                vel[k] = v;
                double x = mid + half * v;
                pos[k] = x;
                double d = x - pos_build[k];
                d2 += d * d;
            }
            moved = check && !(d2 <= lim2);
        }
    }
    wave_raise_flags(flag, moved, si);
}

}  // namespace

MdContext::MdContext(const omm_md_desc* descs, int nsys, int precision) {
    if (!descs || nsys <= 0) throw Error(OMM_ERR_INVALID, "null MD descriptor");
    nsys_ = nsys;
    std::vector<const omm_nb_desc*> nbd(nsys);
    std::vector<int> ns(nsys);
    std::vector<const double*> boxes(nsys);
    for (int r = 0; r < nsys; ++r) {
        const omm_md_desc* d = descs + r;
        if (!d->nonbonded) throw Error(OMM_ERR_INVALID, "null MD descriptor");
        if (d->num_particles <= 0 || !d->masses) throw Error(OMM_ERR_INVALID, "MD descriptor needs particles and masses");
        if (d->nonbonded->num_particles != d->num_particles)
            throw Error(OMM_ERR_INVALID, "NonbondedForce particle count mismatch");
        if (!(std::isfinite(d->temperature) && d->temperature >= 0.0))
            throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
        if (!(std::isfinite(d->friction) && d->friction >= 0.0))
            throw Error(OMM_ERR_INVALID, "friction must be nonnegative and finite");
        if (!(std::isfinite(d->step_size) && d->step_size > 0.0))
            throw Error(OMM_ERR_INVALID, "step size must be positive and finite");
        if (d->temperature != descs[0].temperature || d->friction != descs[0].friction ||
            d->step_size != descs[0].step_size)
            throw Error(OMM_ERR_INVALID, "the replicas of a batch must share the integrator settings");
        nbd[r] = d->nonbonded;
        ns[r] = d->num_particles;
        boxes[r] = d->box;
    }
    mixed_ = precision == OMM_PRECISION_MIXED;
    nb_.reset(new NbOperator(nbd.data(), nsys, precision));
    nb_->initialize(ns.data(), boxes.data());
    s_ = nb_->stream();
    n_ = nb_->total_particles();
    std::vector<double> inv;
    std::vector<int> b1, b2, a1, a2, a3;
    std::vector<double> blen, bk, ath, ak;
    for (int r = 0; r < nsys; ++r) {
        const omm_md_desc* d = descs + r;
        const int off = nb_->atom_offset(r), n = d->num_particles;
        for (int i = 0; i < n; ++i) {
            const double m = d->masses[i];
            if (!(std::isfinite(m) && m >= 0.0)) throw Error(OMM_ERR_INVALID, "masses must be nonnegative");
            masses_.push_back(m);
            inv.push_back(m == 0.0 ? 0.0 : 1.0 / m);
        }
        auto idx = [&](const int32_t* p, int k) {
            if (p[k] < 0 || p[k] >= n) throw Error(OMM_ERR_INVALID, "bonded particle index out of range");
            return off + p[k];
        };
        if (d->num_bonds < 0 || d->num_angles < 0) throw Error(OMM_ERR_INVALID, "negative bonded term count");
        for (int k = 0; k < d->num_bonds; ++k) {
            b1.push_back(idx(d->bond_particle1, k));
            b2.push_back(idx(d->bond_particle2, k));
            blen.push_back(d->bond_length[k]);
            bk.push_back(d->bond_k[k]);
        }
        for (int k = 0; k < d->num_angles; ++k) {
            a1.push_back(idx(d->angle_particle1, k));
            a2.push_back(idx(d->angle_particle2, k));
            a3.push_back(idx(d->angle_particle3, k));
            ath.push_back(d->angle_theta0[k]);
            ak.push_back(d->angle_k[k]);
        }
        nb_->systems()[r].seed = d->seed;
    }
    nb_->upload_systems(s_);
    inv_mass_.upload(inv, s_);
    nbonds_ = (int)b1.size();
    nangles_ = (int)a1.size();
    if (nbonds_ > 0) {
        b1_.upload(b1, s_);
        b2_.upload(b2, s_);
        blen_.upload(blen, s_);
        bk_.upload(bk, s_);
    }
    if (nangles_ > 0) {
        a1_.upload(a1, s_);
        a2_.upload(a2, s_);
        a3_.upload(a3, s_);
        ath_.upload(ath, s_);
        ak_.upload(ak, s_);
    }
    dt_ = descs[0].step_size;
    temperature_ = descs[0].temperature;
    friction_ = descs[0].friction;
    pos_.alloc(3 * (size_t)n_);
    vel_.alloc(3 * (size_t)n_);
    fill_async<double>(pos_.p, 3 * (size_t)n_, 0.0, s_);
    fill_async<double>(vel_.p, 3 * (size_t)n_, 0.0, s_);
    const size_t eslots = (size_t)E_SLOTS * nsys;
    if (mixed_) {
        force_fx_.alloc(3 * (size_t)n_);
        energy_fx_.alloc(eslots);
    } else {
        force_d_.alloc(3 * (size_t)n_);
        energy_d_.alloc(eslots);
    }
    out_.alloc(3 * (size_t)n_);
    status_.alloc(1);
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    step_dev_.alloc(1);
    fill_async<unsigned long long>(step_dev_.p, 1, 0ull, s_);
    snap_pos_.alloc(3 * (size_t)n_);
    snap_vel_.alloc(3 * (size_t)n_);
    snap_step_.alloc(1);
    clear_accumulators(s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

MdContext::~MdContext() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
}

void MdContext::set_positions(int r, const double* pos) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    const int n = nb_->num_particles(r);
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i])) throw Error(OMM_ERR_NONFINITE, "particle positions must be finite");
    OMM_HIP(hipMemcpyAsync(pos_.p + 3 * (size_t)nb_->atom_offset(r), pos, 3 * (size_t)n * sizeof(double),
                           hipMemcpyHostToDevice, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    positions_dirty_ = true;
}

void MdContext::prepare() {
    if (!positions_dirty_) return;
    if (nb_->calibrate(pos_.p, status_.p, s_)) drop_graph();
    nb_->force_rebuild(s_);
    OMM_HIP(hipStreamSynchronize(s_));
    positions_dirty_ = false;
}

void MdContext::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

void MdContext::set_velocities(int r, const double* vel) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    const int n = nb_->num_particles(r);
    OMM_HIP(hipMemcpyAsync(vel_.p + 3 * (size_t)nb_->atom_offset(r), vel, 3 * (size_t)n * sizeof(double),
                           hipMemcpyHostToDevice, s_));
    OMM_HIP(hipStreamSynchronize(s_));
}

void MdContext::launch_bonded(hipStream_t s) {
    const NbSettings& c = nb_->settings();
    const ExceptionDev& ex = nb_->exceptions();
    const int ne = c.include_direct ? ex.count : 0;
    const int total = ne + nbonds_ + nangles_;
    if (total == 0) return;
    const int blocks = (total + 255) / 256;
    if (mixed_)
        hipLaunchKernelGGL(k_bonded_all<long long>, dim3(blocks), dim3(256), 0, s, ne, ex.p1.p, ex.p2.p, ex.qq.p,
                           ex.sig.p, ex.eps.p, nb_->charge_dev(), nb_->exceptions_pbc() ? 1 : 0, nb_->has_pme() ? 1 : 0,
                           nb_->alpha(), nbonds_, b1_.p, b2_.p, blen_.p, bk_.p, nangles_, a1_.p, a2_.p, a3_.p, ath_.p,
                           ak_.p, pos_.p, nb_->atom_sys_dev(), nb_->systems_dev(), force_fx_.p, energy_fx_.p,
                           status_.p);
    else
        hipLaunchKernelGGL(k_bonded_all<double>, dim3(blocks), dim3(256), 0, s, ne, ex.p1.p, ex.p2.p, ex.qq.p,
                           ex.sig.p, ex.eps.p, nb_->charge_dev(), nb_->exceptions_pbc() ? 1 : 0, nb_->has_pme() ? 1 : 0,
                           nb_->alpha(), nbonds_, b1_.p, b2_.p, blen_.p, bk_.p, nangles_, a1_.p, a2_.p, a3_.p, ath_.p,
                           ak_.p, pos_.p, nb_->atom_sys_dev(), nb_->systems_dev(), force_d_.p, energy_d_.p, status_.p);
    OMM_HIP(hipGetLastError());
}

void MdContext::enqueue_forces(hipStream_t s, bool energy, int rebuild) {
    const NbSettings& c = nb_->settings();
    nb_->enqueue(pos_.p, c.include_direct, true, energy, rebuild, false, force_fx_.p, force_d_.p, energy_fx_.p,
                 energy_d_.p, status_.p, s);
    launch_bonded(s);
}

void MdContext::clear_accumulators(hipStream_t s) {
    const size_t eslots = (size_t)E_SLOTS * nsys_;
    if (mixed_) {
        fill_async<long long>(force_fx_.p, 3 * (size_t)n_, 0ll, s);
        fill_async<long long>(energy_fx_.p, eslots, 0ll, s);
    } else {
        fill_async<double>(force_d_.p, 3 * (size_t)n_, 0.0, s);
        fill_async<double>(energy_d_.p, eslots, 0.0, s);
    }
}

// One step.  Invariant: the force accumulator is zero on entry (the previous
// integrator cleared it).  The integrator raises the list's rebuild flag when
// an atom left the skin, and the next step's build kernels act on it.
void MdContext::enqueue_step(hipStream_t s, int rebuild) {
    const int eslots = E_SLOTS * nsys_;
    const int eb = std::min(64, (eslots + 255) / 256);
    if (mixed_)
        hipLaunchKernelGGL(k_step_begin<long long>, dim3(eb), dim3(256), 0, s, step_dev_.p, energy_fx_.p, eslots);
    else
        hipLaunchKernelGGL(k_step_begin<double>, dim3(eb), dim3(256), 0, s, step_dev_.p, energy_d_.p, eslots);
    enqueue_forces(s, false, rebuild);
    const double decay = std::exp(-friction_ * dt_);
    const double variance = BOLTZMANN * temperature_ * -std::expm1(-2.0 * friction_ * dt_);
    const double lim = 0.5 * nb_->skin();
    const int blocks = (n_ + 255) / 256;
    const int check = nb_->settings().include_direct ? 1 : 0;
    if (mixed_)
        hipLaunchKernelGGL(k_lfmiddle<long long>, dim3(blocks), dim3(256), 0, s, n_, pos_.p, vel_.p, force_fx_.p,
                           inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance, step_dev_.p,
                           nb_->pos_build(), lim * lim, check, nb_->list_flag());
    else
        hipLaunchKernelGGL(k_lfmiddle<double>, dim3(blocks), dim3(256), 0, s, n_, pos_.p, vel_.p, force_d_.p,
                           inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance, step_dev_.p,
                           nb_->pos_build(), lim * lim, check, nb_->list_flag());
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

// The captured graph holds GRAPH_STEPS steps and a snapshot of the state
// taken before them.  Each step rebuilds the list only when the previous
// integrator flagged it (the build kernels exit at once otherwise).  A list
// overflow inside the graph restores the snapshot, grows the list and replays
// the steps eagerly (identical random numbers: the RNG is counter-based).
void MdContext::capture_graph() {
    OMM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        const size_t n3 = 3 * (size_t)n_;
        hipLaunchKernelGGL(k_copy2, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s_, pos_.p, snap_pos_.p, vel_.p,
                           snap_vel_.p, n3, step_dev_.p, snap_step_.p);
        for (int k = 0; k < GRAPH_STEPS; ++k) enqueue_step(s_, NbOperator::REBUILD_FLAG);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OMM_HIP(hipStreamEndCapture(s_, &graph_));
    OMM_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
}

void MdContext::set_integrator(double temperature, double friction, double step_size) {
    // langevin.rs:82-101
    if (!(std::isfinite(temperature) && temperature >= 0.0))
        throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
    if (!(std::isfinite(friction) && friction >= 0.0))
        throw Error(OMM_ERR_INVALID, "friction must be nonnegative and finite");
    if (!(std::isfinite(step_size) && step_size > 0.0))
        throw Error(OMM_ERR_INVALID, "step size must be positive and finite");
    if (temperature != temperature_ || friction != friction_ || step_size != dt_) drop_graph();
    temperature_ = temperature;
    friction_ = friction;
    dt_ = step_size;
}

void MdContext::step(int nsteps) {
    if (nsteps < 0) throw Error(OMM_ERR_INVALID, "step count must be nonnegative");
    prepare();
    int done = 0;
    if (mixed_ && nsteps >= GRAPH_STEPS) {
        if (!graph_exec_) capture_graph();
        while (nsteps - done >= GRAPH_STEPS) {
            OMM_HIP(hipGraphLaunch(graph_exec_, s_));
            unsigned st = 0;
            OMM_HIP(hipMemcpyAsync(&st, status_.p, sizeof(unsigned), hipMemcpyDeviceToHost, s_));
            OMM_HIP(hipStreamSynchronize(s_));
            if (st & ST_LIST_OVERFLOW) {
                // Restore the snapshot, grow the list and replay eagerly.
                const size_t bytes = 3 * (size_t)n_ * sizeof(double);
                OMM_HIP(hipMemcpyAsync(pos_.p, snap_pos_.p, bytes, hipMemcpyDeviceToDevice, s_));
                OMM_HIP(hipMemcpyAsync(vel_.p, snap_vel_.p, bytes, hipMemcpyDeviceToDevice, s_));
                OMM_HIP(hipMemcpyAsync(step_dev_.p, snap_step_.p, sizeof(unsigned long long),
                                       hipMemcpyDeviceToDevice, s_));
                clear_accumulators(s_);
                fill_async<unsigned>(status_.p, 1, 0u, s_);
                drop_graph();
                nb_->grow_list();
                nb_->force_rebuild(s_);
                for (int k = 0; k < GRAPH_STEPS; ++k) enqueue_step(s_, NbOperator::REBUILD_FLAG);
                check_status();
                capture_graph();
            }
            done += GRAPH_STEPS;
        }
    }
    // Plain FP64 form, and the remainder: one launch sequence per step.
    for (; done < nsteps; ++done) enqueue_step(s_, NbOperator::REBUILD_FLAG);
    step_count_ += nsteps;
    check_status();
}

void MdContext::evaluate(double* forces, double* energy, bool want_energy) {
    prepare();
    clear_accumulators(s_);
    enqueue_forces(s_, want_energy, NbOperator::REBUILD_FLAG);
    const size_t eslots = (size_t)E_SLOTS * nsys_;
    std::vector<double> e(eslots);
    if (mixed_) {
        std::vector<long long> efx(eslots);
        energy_fx_.download(efx.data(), eslots, s_);
        if (forces) {
            launch_fixed_to_double(force_fx_.p, out_.p, 3 * (size_t)n_, s_);
            out_.download(forces, 3 * (size_t)n_, s_);
        }
        OMM_HIP(hipStreamSynchronize(s_));
        for (size_t k = 0; k < eslots; ++k) e[k] = (double)efx[k] * (1.0 / ENERGY_SCALE);
    } else {
        energy_d_.download(e.data(), eslots, s_);
        if (forces) force_d_.download(forces, 3 * (size_t)n_, s_);
        OMM_HIP(hipStreamSynchronize(s_));
    }
    clear_accumulators(s_);  // the step expects a zero force accumulator
    check_status();
    if (energy) {
        const NbSettings& c = nb_->settings();
        std::vector<SysInfo> sys(nsys_);
        OMM_HIP(hipMemcpyAsync(sys.data(), nb_->systems_dev(), nsys_ * sizeof(SysInfo), hipMemcpyDeviceToHost, s_));
        OMM_HIP(hipStreamSynchronize(s_));
        for (int r = 0; r < nsys_; ++r) {
            const double* er = e.data() + (size_t)E_SLOTS * r;
            HostTerms ht = nb_->host_terms(r, sys[r].box.volume, c.include_direct, true);
            energy[r] = er[E_DIRECT] + er[E_RECIP] + ht.self + ht.dispersion + er[E_EXCEPT] + er[E_BOND] + er[E_ANGLE];
        }
    }
}

void MdContext::energies(double* out) { evaluate(nullptr, out, true); }

void MdContext::set_velocities_to_temperature(double temperature, uint64_t seed) {
    // context.rs:182-223
    if (!(std::isfinite(temperature) && temperature >= 0.0))
        throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
    std::vector<double> v(3 * (size_t)n_), f(3 * (size_t)n_);
    for (int r = 0; r < nsys_; ++r) {
        std::mt19937_64 rng(seed + (uint64_t)r);
        std::normal_distribution<double> normal(0.0, 1.0);
        const int off = nb_->atom_offset(r), n = nb_->num_particles(r);
        for (int i = off; i < off + n; ++i) {
            double scale = masses_[i] == 0.0 ? 0.0 : std::sqrt(BOLTZMANN * temperature / masses_[i]);
            for (int a = 0; a < 3; ++a) v[3 * (size_t)i + a] = normal(rng) * scale;
        }
    }
    evaluate(f.data(), nullptr, false);
    double offset = 0.5 * dt_;  // LangevinMiddle velocity_time_offset (langevin.rs:204-206)
    for (int i = 0; i < n_; ++i) {
        double im = masses_[i] == 0.0 ? 0.0 : 1.0 / masses_[i];
        for (int a = 0; a < 3; ++a) v[3 * (size_t)i + a] -= f[3 * (size_t)i + a] * im * offset;
    }
    vel_.upload(v, s_);
}

void MdContext::get_state(int r, double* pos, double* vel, double* forces, double* energy, int64_t* step) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    const size_t off3 = 3 * (size_t)nb_->atom_offset(r), n3 = 3 * (size_t)nb_->num_particles(r);
    if (forces || energy) {
        std::vector<double> f(forces ? 3 * (size_t)n_ : 0), e(nsys_);
        evaluate(forces ? f.data() : nullptr, energy ? e.data() : nullptr, energy != nullptr);
        if (forces) std::copy(f.begin() + off3, f.begin() + off3 + n3, forces);
        if (energy) *energy = e[r];
    }
    if (pos) OMM_HIP(hipMemcpyAsync(pos, pos_.p + off3, n3 * sizeof(double), hipMemcpyDeviceToHost, s_));
    if (vel) OMM_HIP(hipMemcpyAsync(vel, vel_.p + off3, n3 * sizeof(double), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    if (step) *step = step_count_;
}

void MdContext::get_box(int r, double* box) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    SysInfo S;
    OMM_HIP(hipMemcpyAsync(&S, nb_->systems_dev() + r, sizeof(SysInfo), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    for (int k = 0; k < 9; ++k) box[k] = S.box.v[k];
}

void MdContext::profile(int reps, double ms[8], int64_t* num_pairs) {
    if (reps <= 0) throw Error(OMM_ERR_INVALID, "reps must be positive");
    prepare();
    nb_->profile(pos_.p, reps, force_fx_.p, force_d_.p, energy_fx_.p, energy_d_.p, status_.p, ms, num_pairs, s_);
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
    ms[6] = timed([&] { launch_bonded(s_); });
    // The integrator is timed on scratch copies of the state.
    DevBuf<double> px, vx;
    DevBuf<int> scratch_flag;  // stands in for the list's nsys + 1 rebuild flags
    px.alloc(3 * (size_t)n_);
    vx.alloc(3 * (size_t)n_);
    scratch_flag.alloc((size_t)nsys_ + 1);
    OMM_HIP(hipMemcpyAsync(px.p, pos_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    OMM_HIP(hipMemcpyAsync(vx.p, vel_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    const double decay = std::exp(-friction_ * dt_);
    const double variance = BOLTZMANN * temperature_ * -std::expm1(-2.0 * friction_ * dt_);
    const double lim = 0.5 * nb_->skin();
    ms[7] = timed([&] {
        if (mixed_)
            hipLaunchKernelGGL(k_lfmiddle<long long>, dim3((n_ + 255) / 256), dim3(256), 0, s_, n_, px.p, vx.p,
                               force_fx_.p, inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance,
                               step_dev_.p, nb_->pos_build(), lim * lim, 1, scratch_flag.p);
        else
            hipLaunchKernelGGL(k_lfmiddle<double>, dim3((n_ + 255) / 256), dim3(256), 0, s_, n_, px.p, vx.p,
                               force_d_.p, inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance,
                               step_dev_.p, nb_->pos_build(), lim * lim, 1, scratch_flag.p);
    });
    OMM_HIP(hipStreamSynchronize(s_));
    OMM_HIP(hipEventDestroy(t0));
    OMM_HIP(hipEventDestroy(t1));
    // Profiling accumulated into the step's buffers: restore the invariants
    // (zero forces, zero status) and request a fresh list.
    clear_accumulators(s_);
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    nb_->force_rebuild(s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

double MdContext::kinetic_energy(int r) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    const int off = nb_->atom_offset(r), n = nb_->num_particles(r);
    std::vector<double> v(3 * (size_t)n);
    OMM_HIP(hipMemcpyAsync(v.data(), vel_.p + 3 * (size_t)off, v.size() * sizeof(double), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    double ke = 0.0;
    for (int i = 0; i < n; ++i)
        ke += 0.5 * masses_[off + i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return ke;
}

}  // namespace ommhip
