// MD contexts: the Context/LangevinMiddle step around the nonbonded operator
// for one system or a batch of independent systems (replicas) advanced by the
// same launches.
#pragma once

#include <memory>

#include "nbforce.h"

namespace ommhip {

// Bonded terms of the MD step on the device: exceptions, harmonic bonds and
// harmonic angles, grouped by small molecule (see MdContext::build_bonded).
struct BondedDev {
    int ngroups = 0, ne = 0, nb = 0, na = 0;  // groups; all exceptions / bonds / angles
    int ge = 0, gb = 0, ga = 0;               // of which grouped (stored first)
    DevBuf<int> grp_atoms;                    // [ngroups][4], -1 padded
    DevBuf<int> grp_e, grp_b, grp_a;          // [ngroups + 1] term offsets per group
    DevBuf<int> e1, e2, b1, b2, a1, a2, a3;
    DevBuf<double> eqq, esig, eeps, blen, bk, ath, ak;
};

class MdContext {
public:
    // Replicas share the nonbonded method/cutoff/tolerance and the integrator
    // settings; particles, topology, box and seed are per replica.
    MdContext(const omm_md_desc* descs, int nsys, int precision);
    ~MdContext();
    int nsys() const { return nsys_; }
    int num_particles(int r) const { return nb_->num_particles(r); }
    void set_positions(int r, const double* pos);
    void set_velocities(int r, const double* vel);
    // Replica r draws from seed + r.
    void set_velocities_to_temperature(double temperature, uint64_t seed);
    void step(int nsteps);
    void set_integrator(double temperature, double friction, double step_size);
    // Steps per captured HIP graph in the mixed form (0: launch every step
    // eagerly; the same kernels run either way).
    void set_graph_steps(int steps);
    void get_state(int r, double* pos, double* vel, double* forces, double* energy, int64_t* step);
    // Potential energy of every replica (nsys values).
    void energies(double* out);
    double kinetic_energy(int r);
    void get_box(int r, double* box);
    void list_stats(int64_t* tiles, int64_t* entries) {
        OMM_HIP(hipStreamSynchronize(s_));
        prepare();
        nb_->list_stats(tiles, entries);
    }
    // Per-phase average device times (ms): the six NbOperator::profile
    // phases, [6] bonds + angles, [7] LangevinMiddle update; plus the
    // in-cutoff pair count of the current positions.  State is preserved.
    void profile(int reps, double ms[8], int64_t* num_pairs);

    std::mutex mutex;

private:
    void enqueue_forces(hipStream_t s, bool energy, int rebuild);
    void enqueue_step(hipStream_t s, int rebuild);
    void launch_bonded(hipStream_t s);
    void build_bonded(const std::vector<int>& e1, const std::vector<int>& e2, const std::vector<double>& eqq,
                      const std::vector<double>& esig, const std::vector<double>& eeps, const std::vector<int>& b1,
                      const std::vector<int>& b2, const std::vector<double>& blen, const std::vector<double>& bk,
                      const std::vector<int>& a1, const std::vector<int>& a2, const std::vector<int>& a3,
                      const std::vector<double>& ath, const std::vector<double>& ak);
    void clear_accumulators(hipStream_t s);
    void capture_graph();
    // Evaluates forces (all atoms, FP64 out, nullable) and the potential
    // energy of every replica (nullable).
    void evaluate(double* forces, double* energy, bool want_energy);
    void check_status();
    unsigned read_status();
    void restore_snapshot();
    void drop_graph();
    // Sizes the neighbour list for the current positions once they are all set.
    void prepare();

    std::unique_ptr<NbOperator> nb_;
    int nsys_ = 1;
    bool mixed_ = true;
    int n_ = 0;  // atoms over all replicas
    bool positions_dirty_ = true;
    hipStream_t s_ = nullptr;
    std::vector<double> masses_;
    double dt_ = 0.001, temperature_ = 0.0, friction_ = 0.0;
    int64_t step_count_ = 0;
    int graph_steps_ = 10;
    DevBuf<double> pos_, vel_, inv_mass_, out_;
    BondedDev bonded_;
    DevBuf<long long> force_fx_, energy_fx_;
    DevBuf<double> force_d_, energy_d_;
    DevBuf<unsigned> status_;
    DevBuf<unsigned long long> step_dev_, snap_step_;
    DevBuf<double> snap_pos_, snap_vel_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace ommhip
