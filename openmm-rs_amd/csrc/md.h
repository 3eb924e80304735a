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
    DevBuf<double> ec6;  // LJPME: exception particles' C6
};

// Monte Carlo barostat state of one system (barostat.rs:42-57, 129-195) plus
// the record of its last attempt.
struct BaroState {
    double volume_scale;                      // maximum |dV| of a proposal
    long long attempted, accepted;            // since the last step-size adaptation
    long long total_attempted, total_accepted;
    double u_volume, u_accept;                // the attempt's two uniforms
    double v_old, v_new, e_old, e_new, work;  // last attempt
    int accepted_last, skipped;
    // Anisotropic barostat (anisotropic_barostat.rs): the same per axis, and
    // the axis of the last proposal.
    double axis_volume_scale[3];
    long long axis_attempted[3], axis_accepted[3];
    int axis;
    double scale[3];     // box scale factors of the last proposal
    double area_change;  // membrane barostat: change of the xy area (nm^2)
    double log_shape;    // flexible barostat: 2 ln s_x + ln s_y of the last proposal
    double box_new[9];   // the proposal's box vectors
};

// Distance constraints (constraints.rs) of massive particles, grouped into
// connected components: one thread projects one component, sweeping its
// constraints in the reference's order (constraints only couple atoms of a
// component, so this is the reference's global Gauss-Seidel loop).
struct ConstraintDev {
    int ncomp = 0, ncons = 0;
    double tolerance = 1e-5;
    DevBuf<int> atom_off, atoms;  // [ncomp + 1] prefix, atoms of each component (ascending)
    DevBuf<int> cons_off;         // [ncomp + 1] prefix into the constraint arrays
    DevBuf<int> c1, c2;           // global particle indices, reference order within a component
    DevBuf<double> d2, w1, w2;    // distance^2, inverse masses
    DevBuf<unsigned char> constrained;  // [N]: integrated by the constraint kernel
    DevBuf<double> prev, proposed;      // [3N] step scratch
};

struct BaroDev {
    bool present = false;  // a barostat is set (its molecules are known), perhaps at frequency 0
    bool enabled = false;  // attempts are made (frequency > 0)
    int frequency = 0, phase = 0;  // phase: steps since the last attempt slot
    double pressure = 0.0, temperature = 0.0;
    // 0 isotropic (MonteCarloBarostat), 1 anisotropic (one axis per
    // attempt), 2 membrane (MonteCarloMembraneBarostat: xy coupled or not,
    // z free / fixed / at constant volume, surface tension), 3 flexible
    // (MonteCarloFlexibleBarostat: every lower-triangle box element moves;
    // volume_scale then holds the length step).
    int kind = 0;
    double axis_pressure[3] = {};
    int scale_axes = 7;            // anisotropic: bit d lets axis d be scaled
    int xy_mode = 0, z_mode = 0;   // membrane: xy 0 isotropic / 1 anisotropic; z 0 free / 1 fixed / 2 constant volume
    double surface_tension = 0.0;  // membrane: bar nm
    uint64_t seed = 0;
    int nmol = 0;
    DevBuf<int> mol_start, mol_atoms, mol_sys, nmol_sys;
    std::vector<int> mol_start_h, mol_atoms_h, mol_sys_h;
    DevBuf<double> transform, trial_box;  // pressure queries: [9 nsys] each
    DevBuf<BaroState> st, snap_st;
    DevBuf<Box> saved_box;
    DevBuf<int> rejected;
    DevBuf<double> scale, pos_saved;
    DevBuf<long long> force_fx, e_old_fx, e_new_fx;
    DevBuf<double> force_d, e_old_d, e_new_d;
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
    // Context::set_parameter (context.rs:251-261) for the NonbondedForce's
    // global parameters; returns whether the force has the parameter.
    bool set_parameter(const std::string& name, double value);
    // MonteCarloBarostat (barostat.rs:59-218) on every replica: isotropic
    // moves every `frequency` steps (0: none, the barostat stays set for the
    // pressure queries), molecules scaled as rigid bodies, replica r drawing
    // from seed + r.  alpha and the PME grid stay
    // as initialised, as in the reference.
    void set_barostat(double pressure, double temperature, int frequency, uint64_t seed);
    // MonteCarloAnisotropicBarostat (anisotropic_barostat.rs:42-243): per-axis
    // pressures (bar); each attempt scales one enabled axis (scale_axes bit d).
    void set_anisotropic_barostat(const double pressure[3], double temperature, int scale_axes, int frequency,
                                  uint64_t seed);
    // MonteCarloFlexibleBarostat (flexible_barostat.rs:132-238).
    void set_flexible_barostat(double pressure, double temperature, int frequency, uint64_t seed);
    // MonteCarloMembraneBarostat (membrane_barostat.rs:167-292).
    void set_membrane_barostat(double pressure, double surface_tension, double temperature, int xy_mode, int z_mode,
                               int frequency, uint64_t seed);
    // Pressure queries of the barostat set on the replicas (context/pressure.rs),
    // per replica: central differences of the FP64 potential energy over
    // trial boxes deformed by +-0.1 % with the molecules moved rigidly, plus
    // the kinetic term of the molecular momenta.  Positions, velocities, boxes
    // and the step count are unchanged.
    void current_pressure(double* out);                     // MonteCarloBarostat: 1 value per replica
    void axis_pressures(int kind, double* out);             // anisotropic (1) / membrane (2): 3 per replica
    void flexible_pressure(double* out);                    // flexible: 6 per replica (xx yy zz xy xz yz)
    void stress_tensor(bool include_kinetic, double* out);  // flexible: 6 per replica
    // One barostat attempt now, outside the step sequence (test hook).
    void barostat_attempt();
    BaroState barostat_state(int r);
    void get_state(int r, double* pos, double* vel, double* forces, double* energy, int64_t* step);
    // Context::set_state (context.rs:274-293) for replica r: the step count
    // (shared by the replicas of a batch: it keys every replica's noise), box
    // vectors, positions and velocities, each optional (null: kept).  The
    // neighbour list of replica r is rebuilt from list_pos when given (the
    // positions the saved context last built its list from, list_positions()),
    // else from pos, so a context restored with its list positions continues
    // the saved trajectory bit for bit.
    void set_state(int r, int64_t step, const double* box, const double* pos, const double* vel,
                   const double* list_pos);
    // Positions replica r's current neighbour list was built from.
    void list_positions(int r, double* out);
    // Test hook of the overflow recovery: list and spread-bin capacities (0:
    // unchanged), see NbOperator::set_capacity; the lists are rebuilt from the
    // positions they were built from.
    void set_capacity(int list_cap, int bin_cap);
    // Current capacities and the number of runs of steps replayed after an
    // overflow since the context was created.
    void capacity(int* list_cap, int* bin_cap, int64_t* replays);
    // Potential energy of every replica (nsys values).
    void energies(double* out);
    double kinetic_energy(int r);
    void get_box(int r, double* box);
    void list_stats(int64_t* tiles, int64_t* entries) {
        OMM_HIP(hipStreamSynchronize(s_));
        prepare();
        nb_->list_stats(tiles, entries);
    }
    // The autotuner's choices (made by the first preparation).
    NbOperator::Tuning tuning() {
        prepare();
        return nb_->tuning();
    }
    // Non-periodic methods: chooses the virtual box again when the current
    // positions (and, when given, a second configuration on the device) do
    // not fit it, and centres the integrators' region on them; true when the
    // device was set up afresh (its lists must be built).
    bool refit_virtual_box(const double* extra_dev);
    void require_periodic_for_barostat() const;
    // Launch shapes set explicitly (NbOperator::apply_tuning); captured
    // graphs are rebuilt.  Results do not change.
    void set_tuning(const NbOperator::Tuning& t);
    // Per-phase average device times (ms): the six NbOperator::profile
    // phases, [6] bonds + angles, [7] LangevinMiddle update; plus the
    // in-cutoff pair count of the current positions.  State is preserved.
    void profile(int reps, double ms[8], int64_t* num_pairs);

    std::mutex mutex;

    // Context::apply_constraints / apply_velocity_constraints (context.rs:767-810):
    // projects the positions (against themselves) or the velocities (against
    // the positions) onto the constraints with this tolerance.
    void project_constraints(double tolerance, bool velocities);

private:
    double default_box_[9] = {};  // a non-periodic system's default box vectors
    void enqueue_forces(hipStream_t s, bool energy, int rebuild);
    void enqueue_forces_into(hipStream_t s, bool energy, int rebuild, long long* force_fx, double* force_d,
                             long long* energy_fx, double* energy_d);
    // One step; its integrator draws at step counter + offset, and `advance`
    // > 0 ends a run of that many steps (advances the counter, clears the
    // energies).  Eager steps use (1, 1); a captured graph (k + 1, 0) and
    // (len, len) for its last step.
    void enqueue_step(hipStream_t s, int rebuild, bool barostat, int offset = 1, int advance = 1);
    void enqueue_barostat(hipStream_t s);
    void require_barostat(int kind, const char* what) const;
    // Context::energy_at over the replicas (context.rs:942-1029): replica r's
    // molecules moved rigidly by transforms[9 r ..] (row-major), its box set to
    // boxes[9 r ..], the potential energies evaluated, and positions and boxes
    // restored bit for bit.
    void energies_at(const std::vector<double>& transforms, const std::vector<double>& boxes, double* energies);
    // pressure.rs:107-131 with rigid molecules: 1/2 P_i P_j / M summed over the
    // molecules of mass M > 0, 6 components per replica.
    void kinetic_tensor(std::vector<double>& out);
    void setup_constraints(const omm_md_desc* descs, int nsys, const std::vector<double>& inv);
    void launch_bonded(hipStream_t s, long long* force_fx, double* force_d, long long* energy_fx, double* energy_d);
    // A barostat period of at most MAX_PERIOD_GRAPH steps is captured whole,
    // its attempt opening the last step; longer periods replay attempt-free
    // graphs of graph_steps_ steps and launch the attempt steps eagerly, so a
    // graph never holds more than max(graph_steps_, MAX_PERIOD_GRAPH) steps.
    static constexpr int MAX_PERIOD_GRAPH = 1000;
    bool period_graph() const { return baro_.enabled && baro_.frequency <= MAX_PERIOD_GRAPH; }
    int graph_length() const { return period_graph() ? baro_.frequency : graph_steps_; }
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
    void enqueue_snapshot(hipStream_t s);
    void restore_snapshot();
    void run_steps(int count, bool graph);
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
    // Positions each system's neighbour list is built from at the next
    // preparation, and which systems were set since the last one (the others
    // rebuild from the positions of their current list, which reproduces it).
    DevBuf<double> list_src_;
    std::vector<char> touched_;
    BondedDev bonded_;
    ConstraintDev cons_;
    std::vector<int> bonded_exc_order_;  // exception k of the bonded arrays = global exception order[k]
    BaroDev baro_;
    std::vector<int> bond1_h_, bond2_h_;  // bonds in global indices (molecules)
    std::vector<int> cons1_h_, cons2_h_;  // constraints in global indices (molecules)
    DevBuf<SysInfo> snap_sys_;
    DevBuf<long long> force_fx_, energy_fx_;
    DevBuf<double> force_d_, energy_d_;
    DevBuf<unsigned> status_;
    DevBuf<unsigned long long> step_dev_, snap_step_;
    DevBuf<double> snap_pos_, snap_vel_, snap_list_;
    int64_t overflow_replays_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace ommhip
