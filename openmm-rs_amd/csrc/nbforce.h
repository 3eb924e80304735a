// NonbondedForce operator on one gfx950 device: the host object behind the
// omm_nb_* entry points and the nonbonded part of the MD context.
#pragma once

#include <mutex>

#include "pme.h"

namespace ommhip {

struct NbSettings {
    std::vector<double> charge, sigma, epsilon;
    std::vector<int> exc1, exc2;
    std::vector<double> exc_qq, exc_sig, exc_eps;
    int method = OMM_NB_NO_CUTOFF;
    double cutoff = 1.0, rf_dielectric = 78.3, tolerance = 5e-4;
    double pme_alpha = 0.0;
    int pme_grid[3] = {0, 0, 0};
    bool dispersion_correction = true, exceptions_pbc = false, switching = false;
    double switching_distance = -1.0;
    int force_group = 0, reciprocal_group = -1;
    bool include_direct = true;

    bool periodic() const {
        return method == OMM_NB_CUTOFF_PERIODIC || method == OMM_NB_EWALD || method == OMM_NB_PME ||
               method == OMM_NB_LJPME;
    }
};

// Validates a descriptor the way the reference's builders do and copies it.
NbSettings settings_from_desc(const omm_nb_desc* d);

// Energy terms computed on the host in FP64 for a given volume.
struct HostTerms {
    double self = 0.0;        // -k (alpha/sqrt(pi) sum q^2 + pi Q^2 / (2 alpha^2 V))
    double dispersion = 0.0;  // dispersion correction
};

class NbOperator {
public:
    NbOperator(const omm_nb_desc* desc, int precision);
    ~NbOperator();

    // Force::initialize (nonbonded.rs:1055-1107).
    void initialize(int n, const double* box);
    bool initialized() const { return initialized_; }
    double alpha() const { return alpha_; }
    const int* grid() const { return grid_; }
    const NbSettings& settings() const { return cfg_; }
    bool mixed() const { return mixed_; }
    int num_particles() const { return n_; }

    // Host-pointer evaluation (calc_forces_and_energy[_groups]).  parts
    // receives the energy split; stats the pair count and hash.
    void compute(const double* pos, const double* box, bool direct, bool recip, double* forces, double* energy,
                 double parts[5], int64_t* num_pairs, uint64_t* pair_hash);

    // Device-resident enqueue for the MD step: neighbour-list check/rebuild,
    // direct pairs, exceptions and PME into the caller's accumulators.  The
    // box lives on the device at box_dev.  Graph-capturable.
    // energy: also accumulate energies (slower, accurate path in the mixed form).
    void enqueue(const double* pos_dev, const Box* box_dev, bool direct, bool recip, bool energy, long long* force_fx,
                 double* force_d, long long* energy_fx, double* energy_d, unsigned* status, hipStream_t s);
    // PME arithmetic for a force-only or an energy evaluation.
    int mesh_mode(bool energy) const;

    // Average device time (ms) of each phase over `reps` launches, measured
    // with HIP events on stream s: [0] neighbour-list rebuild (forced),
    // [1] direct pairs, [2] exceptions, [3] PME spread, [4] PME FFT passes +
    // convolution, [5] PME gather.  Accumulates into the given buffers.
    void profile(const double* pos_dev, const Box* box_dev, int reps, long long* force_fx, double* force_d,
                 long long* energy_fx, double* energy_d, unsigned* status, double ms[6], int64_t* num_pairs,
                 hipStream_t s);

    HostTerms host_terms(double volume, bool direct, bool recip) const;
    void set_stream(hipStream_t s) { stream_ = s; }
    hipStream_t stream() const { return stream_; }
    // Rebuild capacity after a list overflow.
    void grow_list();
    // Builds the list for these positions and sizes its capacity to twice the
    // longest j-list (host-side, outside graph capture).  Returns true when the
    // list storage was reallocated (captured graphs must then be rebuilt).
    bool calibrate(const double* pos_dev, const Box* box_dev, unsigned* status, hipStream_t s);
    void force_rebuild(hipStream_t s);

    std::mutex mutex;

private:
    void check_box(const double* box) const;
    double dispersion_coefficient() const;

    NbSettings cfg_;
    bool mixed_ = true;
    bool initialized_ = false;
    int n_ = 0;
    double alpha_ = 0.0;
    int grid_[3] = {0, 0, 0};
    double sum_q2_ = 0.0, sum_q_ = 0.0;
    mutable double disp_coeff_ = 0.0;
    mutable bool disp_cached_ = false;
    int pair_grid_ = 1;
    hipStream_t stream_ = nullptr;
    double box0_[9] = {0};

    ParticleDev part_;
    ExceptionDev exc_;
    NlistDev nl_;
    PmeDev pme_;
    PairParams pp_;

    // Host-API scratch.
    DevBuf<double> pos_dev_;
    DevBuf<Box> box_dev_;
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
