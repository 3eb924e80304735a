// MD context: the Context/LangevinMiddle step around the nonbonded operator.
#pragma once

#include <memory>

#include "nbforce.h"

namespace ommhip {

class MdContext {
public:
    MdContext(const omm_md_desc* d, int precision);
    ~MdContext();
    void set_positions(const double* pos);
    void set_velocities(const double* vel);
    void set_velocities_to_temperature(double temperature, uint64_t seed);
    void step(int nsteps);
    void get_state(double* pos, double* vel, double* forces, double* energy, int64_t* step);
    double kinetic_energy();
    // Per-phase average device times (ms): the six NbOperator::profile
    // phases, [6] bonds + angles, [7] LangevinMiddle update; plus the
    // in-cutoff pair count of the current positions.  State is preserved.
    void profile(int reps, double ms[8], int64_t* num_pairs);

    std::mutex mutex;

private:
    void enqueue_forces(hipStream_t s, bool energy);
    void enqueue_step(hipStream_t s);
    double evaluate(double* forces, bool energy);
    void check_status();
    void drop_graph();

    std::unique_ptr<NbOperator> nb_;
    bool mixed_ = true;
    int n_ = 0;
    int nbonds_ = 0, nangles_ = 0;
    hipStream_t s_ = nullptr;
    std::vector<double> masses_;
    Box box_h_{};
    double dt_ = 0.001, temperature_ = 0.0, friction_ = 0.0;
    uint64_t seed_ = 0;
    int64_t step_count_ = 0;
    DevBuf<double> pos_, vel_, inv_mass_, out_;
    DevBuf<int> b1_, b2_, a1_, a2_, a3_;
    DevBuf<double> blen_, bk_, ath_, ak_;
    DevBuf<Box> box_;
    DevBuf<long long> force_fx_, energy_fx_;
    DevBuf<double> force_d_, energy_d_;
    DevBuf<unsigned> status_;
    DevBuf<unsigned long long> step_dev_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace ommhip
