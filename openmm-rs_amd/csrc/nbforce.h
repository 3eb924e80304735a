// NonbondedForce operator on one gfx950 device: the host object behind the
// omm_nb_* entry points (one system) and the nonbonded part of the MD
// contexts (one or many systems evaluated by the same launches).
#pragma once

#include <mutex>
#include <string>

#include "ewald.h"
#include "pme.h"

namespace ommhip {

struct ParamOffset {
    int parameter = 0, index = 0;
    double scale[3] = {0, 0, 0};  // charge (product), sigma, epsilon
};

struct NbSettings {
    // Effective parameters (base + offsets at the current global values).
    std::vector<double> charge, sigma, epsilon;
    std::vector<int> exc1, exc2;
    std::vector<double> exc_qq, exc_sig, exc_eps;
    // Base parameters, global parameters and offsets (nonbonded.rs:247-430).
    std::vector<double> base_charge, base_sigma, base_epsilon, base_qq, base_sig, base_eps;
    std::vector<std::string> gp_names;
    std::vector<double> gp_defaults, gp_values;
    std::vector<ParamOffset> particle_offsets, exception_offsets;
    // Recomputes the effective parameters (nonbonded.rs:395-430).
    void apply_offsets();
    int method = OMM_NB_NO_CUTOFF;
    double cutoff = 1.0, rf_dielectric = 78.3, tolerance = 5e-4;
    double pme_alpha = 0.0;
    int pme_grid[3] = {0, 0, 0};
    double ljpme_alpha = 0.0;  // 0: derived from the tolerance
    int ljpme_grid[3] = {0, 0, 0};
    bool dispersion_correction = true, exceptions_pbc = false, switching = false;
    double switching_distance = -1.0;
    int force_group = 0, reciprocal_group = -1;
    bool include_direct = true;

    bool periodic() const {
        return method == OMM_NB_CUTOFF_PERIODIC || method == OMM_NB_EWALD || method == OMM_NB_PME ||
               method == OMM_NB_LJPME;
    }
    bool use_cutoff() const { return method != OMM_NB_NO_CUTOFF; }
    // Ewald-screened direct space with a reciprocal part (PME, LJPME or plain Ewald).
    bool screened() const { return method == OMM_NB_PME || method == OMM_NB_EWALD || method == OMM_NB_LJPME; }
    // A smooth-PME Coulomb mesh (PME, LJPME).
    bool pme_mesh() const { return method == OMM_NB_PME || method == OMM_NB_LJPME; }
    bool reaction_field() const {
        return method == OMM_NB_CUTOFF_PERIODIC || method == OMM_NB_CUTOFF_NON_PERIODIC;
    }
};

// C6 = 4 sqrt(eps_i eps_j) (sigma_i sigma_j)^3 of every exception's particles
// (nonbonded.rs:1022-1027) under LJPME, zeros otherwise.
std::vector<double> exception_c6(const NbSettings& c);

// Validates a descriptor the way the reference's builders do and copies it.
NbSettings settings_from_desc(const omm_nb_desc* d);

// Energy terms computed on the host in FP64 for a given volume.
struct HostTerms {
    double self = 0.0;        // -k (alpha/sqrt(pi) sum q^2 + pi Q^2 / (2 alpha^2 V))
    double dispersion = 0.0;  // dispersion correction
};

// Path of the process-wide tuning cache (NULL or empty: none); replaces the
// OMMHIP_TUNING_CACHE environment variable.
void set_tuning_cache_path(const char* path);

class NbOperator {
public:
    // A batch of nsys systems.  Method, cutoff, tolerance, dielectric,
    // dispersion correction, exception PBC, switching, force groups and
    // direct-space flag must agree across the batch; particle and exception
    // parameters, sizes and boxes are per system.
    NbOperator(const omm_nb_desc* const* descs, int nsys, int precision);
    ~NbOperator();

    // Force::set_global_parameter (nonbonded.rs:1113-1123) on every system
    // that has the parameter; returns whether any has it.  Changed values
    // refresh the device parameters (and *changed is set).
    bool set_global_parameter(const std::string& name, double value, bool* changed = nullptr);
    bool get_global_parameter(const std::string& name, double* value) const;
    // NonbondedForce::update_parameters_from (nonbonded.rs:721-767) of a
    // one-system operator.
    void update_parameters(const omm_nb_desc* desc);

    // Force::initialize (nonbonded.rs:1055-1107) of every system.
    void initialize(const int* n, const double* const* boxes);
    void initialize(int n, const double* box) { initialize(&n, &box); }
    bool initialized() const { return initialized_; }
    int nsys() const { return (int)cfg_.size(); }
    double alpha(int s = 0) const { return alpha_[s]; }
    const int* grid(int s = 0) const { return grid_[s].data(); }
    double lj_alpha(int s = 0) const { return lj_alpha_[s]; }
    const int* lj_grid(int s = 0) const { return lj_grid_[s].data(); }
    bool ljpme() const { return cfg_[0].method == OMM_NB_LJPME; }
    const NbSettings& settings(int s = 0) const { return cfg_[s]; }
    // Periodic box checks of every evaluation (nonbonded.rs:802-812).
    void check_box(const double* box) const;
    bool mixed() const { return mixed_; }
    int num_particles(int s = 0) const { return n_[s]; }
    int total_particles() const { return ntotal_; }
    int atom_offset(int s) const { return sys_h_[s].atom_off; }

    // Host-pointer evaluation of a one-system operator (calc_forces_and_energy
    // [_groups]).  parts receives the energy split; stats the pair count and
    // hash.
    // Device form: pos_dev_in (device N x 3, read after the caller's stream
    // reaches the call) replaces pos, and forces_dev_out (device N x 3,
    // overwritten) receives the forces; the call returns with them complete.
    void compute(const double* pos, const double* box, bool direct, bool recip, double* forces, double* energy,
                 double parts[5], int64_t* num_pairs, uint64_t* pair_hash, const double* pos_dev_in = nullptr,
                 double* forces_dev_out = nullptr, hipStream_t caller = nullptr);

    // Neighbour-list policy of one enqueued evaluation.
    enum Rebuild : int {
        REBUILD_CHECK = 0,   // displacement check, then a rebuild only if needed
        REBUILD_FORCED = 1,  // rebuild unconditionally
        REBUILD_NONE = 2,    // the caller knows the list is fresh
        REBUILD_FLAG = 3     // rebuild if the list's flag is set (raised by the integrator)
    };
    // Device-resident enqueue for the MD step: neighbour-list rebuild per the
    // policy, direct pairs, exceptions (unless the caller evaluates them
    // itself) and PME into the caller's accumulators (forces over all atoms,
    // energies E_SLOTS per system).  Graph-capturable.  energy: also
    // accumulate energies (slower, accurate path in the mixed form).
    // The PME mesh runs on a side stream concurrently with the pairs; with
    // join_now false the caller enqueues more work on s and then calls join(s).
    void enqueue(const double* pos_dev, bool direct, bool recip, bool energy, int rebuild, bool exceptions,
                 long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
                 hipStream_t s, bool join_now = true);
    void join(hipStream_t s);
    // PME arithmetic for a force-only or an energy evaluation.
    int mesh_mode(bool energy) const;

    // Device data the MD step's fused kernels use.
    const ExceptionDev& exceptions() const { return exc_; }
    const double* charge_dev() const { return part_.charge.p; }
    const int* atom_sys_dev() const { return part_.atom_sys.p; }
    bool exceptions_pbc() const { return cfg_[0].exceptions_pbc; }
    // Non-periodic method on a virtual box (one system).
    bool is_virtual() const { return virtual_; }
    // MD contexts of a non-periodic method: keeps the virtual box while the
    // `count` host positions (any number of configurations of the system,
    // concatenated) fit it, else chooses a new one and sets the device up for
    // it.  Writes the region the integrators check.  Returns true when the
    // device was set up (lists must be rebuilt, captured graphs dropped).
    bool fit_virtual_box(const double* pos, int count);
    // Exceptions carry the reciprocal-space correction (PME or Ewald).
    bool screened() const { return cfg_[0].screened(); }
    const double* pos_build() const { return nl_.pos_build.p; }
    int* list_flag() { return nl_.flag.p; }
    double skin() const { return nl_.skin; }
    // System table: host copy and device table.  Boxes changed on the host
    // are pushed with upload_systems().
    std::vector<SysInfo>& systems() { return sys_h_; }
    SysInfo* systems_dev() { return sys_dev_.p; }
    void upload_systems(hipStream_t s);

    // Average device time (ms) of each phase over `reps` launches, measured
    // with HIP events on stream s: [0] neighbour-list rebuild (forced),
    // [1] direct pairs, [2] exceptions, [3] PME spread, [4] PME FFT passes +
    // convolution, [5] PME gather.  Accumulates into the given buffers.
    void profile(const double* pos_dev, int reps, long long* force_fx, double* force_d, long long* energy_fx,
                 double* energy_d, unsigned* status, double ms[6], int64_t* num_pairs, hipStream_t s);

    HostTerms host_terms(int s, double volume, bool direct, bool recip) const;
    hipStream_t stream() const { return stream_; }
    // Grows the neighbour-list capacity (ST_LIST_OVERFLOW) and/or the PME
    // spread bins (ST_BIN_OVERFLOW) named by the status bits.
    void grow(unsigned status);
    // Builds the list for these positions and sizes its capacity to twice the
    // longest j-list (host-side, outside graph capture).  Returns true when the
    // list storage was reallocated (captured graphs must then be rebuilt).
    bool calibrate(const double* pos_dev, unsigned* status, hipStream_t s);
    // The meshes' own slot order from the current list (after a list build
    // outside the step loop).
    void take_mesh_order(hipStream_t s);
    // Every mesh holds a slot order of the current list layout: the whole
    // mesh, gather included, may then run beside a list build.
    bool mesh_order_ready() const;
    // Launch-shape autotuner, run once by the first calibration: the spread
    // tile edge, the FFT decomposition and lines per workgroup, and the pair
    // kernel's grid, each picked by timing the candidates on the current
    // positions, or taken from the tuning cache when it holds this device
    // and workload.  None of them changes a result bit.
    void autotune(const double* pos_dev, unsigned* status, hipStream_t s);
    struct Tuning {
        int spread_edge = 0, fft_cap = 0, pair_grid = 0;
        int fft_plane = 0;   // 1: plane passes around the x pass, 0: five line passes
        int pair_scale = 1;  // pair grid per calibrated task count: 1, 2 (doubled) or -2 (halved)
        double spread_ms = 0, fft_ms = 0, pair_ms = 0;
        int from_cache = 0;  // taken from the tuning cache instead of timed
    };
    const Tuning& tuning() const { return tuning_; }
    // Applies launch shapes (spread edge, FFT decomposition and lines, pair
    // scale) without timing: the tuning cache's replay and a test hook.
    // Throws OMM_ERR_INVALID for values outside the candidates, or for plane
    // passes that do not fit this workload.
    void apply_tuning(const Tuning& t, const double* pos_dev, unsigned* status, hipStream_t s);
    void force_rebuild(hipStream_t s);
    // Rebuild requests recomputed from these positions: cleared, then raised
    // for every system with an atom more than skin/2 from its position at
    // the last build (the integrators' test).
    void recheck_list(const double* pos_dev, hipStream_t s);
    // Box of system s in the device table and the host copy, checked as at
    // initialize (reduced form, cutoff at most half of each edge; Ewald:
    // rectangular).  Set-state path (Context::set_periodic_box_vectors).
    void set_box(int s, const double* box9, hipStream_t st);
    // Test hook: reallocates the list storage to `list_cap` entries per
    // cluster (0: unchanged; rounded up to whole tiles) and rebuilds the lists
    // from src, and the PME spread bins to `bin_cap` atoms per tile (0:
    // unchanged).  Capacities below the need raise the overflow status (at
    // once for the lists, at the next spread for the bins); every write stays
    // within the storage.
    void set_capacity(int list_cap, int bin_cap, const double* src, unsigned* status, hipStream_t s);
    // PmeCache::evaluate_charges / charge_derivatives (pme.rs:63-138) on a
    // one-system plain-FP64 operator dedicated to the charge mesh: the
    // Coulomb mesh of per-call charges at these positions, in this box, with
    // this alpha (the grid is the operator's).  With count == 0 the mesh
    // forces are added to forces (nullable) and the mesh energy is returned;
    // with count > 0 the mesh potential at the stencil of each listed atom
    // (dE/dq) is added to derivatives[k] instead.  The box and alpha stay in
    // the operator's system table for the next call.
    double mesh_charges(const double* pos, const double* box, double alpha, const double* charges, double* forces,
                        const int* atoms, int count, double* derivatives);
    // Current list capacity (entries per cluster) and spread-bin capacity.
    void capacity(int* list_cap, int* bin_cap) const {
        *list_cap = nl_.cap;
        *bin_cap = pme_.bin_cap;
    }
    // Tile count and total j-list length of the current list (host read-back).
    void list_stats(int64_t* tiles, int64_t* entries);

    std::mutex mutex;

private:
    // Uploads the effective particle and exception parameters and the
    // parameter-dependent energy constants of every system.
    void refresh_parameters();
    // Copies the launch layout fields of sys_h_ (spread tiles, FFT lines)
    // into the device table, keeping its boxes.
    void patch_layout(hipStream_t s);
    // Key of this device and workload in the tuning cache.
    std::string tuning_key() const;
    // Pair grid for a calibrated task count and the tuned scale.
    static int pair_grid_for(long long tasks, int scale);
    // Grows the spread bins until the positions fit.
    void fit_bins(const double* pos_dev, unsigned* status, hipStream_t s);
    void upload_parameters(hipStream_t s);
    // Device set-up of every system for these (validated) boxes: list,
    // mesh, kernel parameters and buffers.
    void setup_device(const std::vector<Box>& boxes);
    // Non-periodic methods (NoCutoff, CutoffNonPeriodic) run the periodic
    // machinery on a virtual rectangular box: one large enough that the
    // minimum image never changes a displacement the list or the kernels
    // look at, and never brings a far pair within reach.  Returns the box
    // for these positions, keeping the current one while it still fits.
    Box virtual_box_for(const double* pos, int count, bool* changed);
    double dispersion_coefficient(int s) const;
    // LJPME self term alpha^6 / 3 sum_i eps_i sigma_i^6 (nonbonded.rs:984-988).
    double lj_self_energy(int s) const;

    std::vector<NbSettings> cfg_;
    bool mixed_ = true;
    bool initialized_ = false;
    bool device_ready_ = false;  // setup_device has run (deferred for non-periodic methods)
    bool virtual_ = false;       // non-periodic method on a virtual box
    Box vbox_{};
    std::vector<int> n_;
    int ntotal_ = 0;
    std::vector<double> alpha_;
    std::vector<std::array<int, 3>> grid_;
    std::vector<double> lj_alpha_;  // LJPME dispersion mesh
    std::vector<std::array<int, 3>> lj_grid_;
    std::vector<double> sum_q2_, sum_q_;
    mutable std::vector<double> disp_coeff_;
    mutable std::vector<char> disp_cached_;
    int pair_grid_ = 1;
    long long calibrated_tasks_ = 0;  // pair-kernel tasks of the calibrated list
    bool tuned_ = false;
    Tuning tuning_;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;  // PME branch (default priority, like stream_)
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    hipEvent_t ev_list_ = nullptr;  // the list build is done (the gather may read the sorted order)
    hipEvent_t ev_caller_ = nullptr;  // device-pointer evaluations: the caller's stream reached the call
    bool joined_ = true;
    bool follow_order_ = false;  // join() copies the list's order into the meshes' own

    std::vector<SysInfo> sys_h_;
    DevBuf<SysInfo> sys_dev_;
    ParticleDev part_;
    ExceptionDev exc_;
    NlistDev nl_;
    PmeDev pme_;
    EwaldDev ewald_;
    // LJPME: the dispersion mesh and its system table (the Coulomb table's
    // layout with the dispersion grid, alpha and influence function; boxes
    // copied from the Coulomb table before each use).
    PmeDev pme_lj_;
    std::vector<SysInfo> sys_lj_h_;
    DevBuf<SysInfo> sys_lj_dev_;
    void lj_mesh(int mode, const double* pos_dev, long long* force_fx, double* force_d, long long* energy_fx,
                 double* energy_d, unsigned* status, hipStream_t s, int phases = PME_ALL, bool atom_order = false);
    PairParams pp_;

    // Host-API scratch.
    DevBuf<double> pos_dev_;
    DevBuf<double> mesh_q_, mesh_out_;  // mesh_charges: charges, potentials
    DevBuf<int> mesh_atoms_;
    DevBuf<long long> force_fx_, energy_fx_;
    DevBuf<double> force_d_, energy_d_;
    DevBuf<unsigned> status_;
    DevBuf<unsigned long long> stats_;
    DevBuf<double> out_forces_;
    Box last_box_{};
    bool have_box_ = false;
};

// Converts int64 fixed-point forces to FP64 (n3 values).
void launch_fixed_to_double(const long long* fx, double* out, size_t n3, hipStream_t s);

}  // namespace ommhip
