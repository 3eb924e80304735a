// Direct-space pair kernel and exception kernel.
#pragma once

#include "nlist.h"

namespace ommhip {

enum ElecKind : int { ELEC_EWALD = 0, ELEC_REACTION_FIELD = 1, ELEC_COULOMB = 2 };
constexpr int PAIR_SWITCH = 8;  // kernel variant flag: LJ switching function on
constexpr int PAIR_LJPME = 16;  // kernel variant flag: LJPME dispersion correction (nonbonded.rs:875-886)
constexpr int ELEC_MASK = 7;

struct PairParams {
    int nblocks = 0, cap = 0;
    int nslots = 0;  // stride of the slot-ordered force components
    int elec = ELEC_EWALD;
    int use_cutoff = 1;
    int periodic = 1;
    double alpha = 0.0;
    double cutoff = 0.0;
    double krf = 0.0, crf = 0.0;  // reaction field (nonbonded.rs:1233-1250)
    int switching = 0;            // LJ switching function (nonbonded.rs:887-899)
    double switch_distance = 0.0;
    double inv_switch_width = 0.0;  // 1 / (cutoff - switch_distance)
    int ljpme = 0;                  // LJPME direct-space dispersion term
    double lj_alpha = 0.0;
};

// Per-particle parameters in the form the kernels use.
struct ParticleDev {
    DevBuf<int> atom_sys;   // system of each atom
    DevBuf<double> charge;  // q (FP64: exceptions, PME)
    DevBuf<float> qk;       // q * sqrt(ONE_4PI_EPS0)
    DevBuf<float> sh;       // sigma / 2
    DevBuf<float> e2;       // 2 sqrt(epsilon)
    DevBuf<double> qk_d, sh_d, e2_d;  // FP64 copies for the plain form
    DevBuf<double> c6amp;             // LJPME mesh amplitude 2 sqrt(epsilon) sigma^3 (pme.rs:50-51)
};

struct ExceptionDev {
    int count = 0;
    DevBuf<int> p1, p2;
    DevBuf<double> qq, sig, eps;
    DevBuf<double> c6;  // LJPME: the particles' C6 (the dispersion correction), else empty
};

// Direct-space pairs over the tile list.  mixed: FP32 math with int64
// fixed-point force/energy accumulation; otherwise FP64 math with FP64 atomics
// into force_d/energy_d.  The pair kernel accumulates into the list's
// slot-ordered buffer; a second kernel adds that into the caller's forces
// (atom order) and clears it.  energy_on: also accumulate the energy (in the mixed
// form each pair energy is evaluated in FP64).  stats (nullable): [0] pair
// count, [1] pair hash per system (no energy is accumulated then).  Energies
// go to slot E_DIRECT of each system's E_SLOTS block.  order_a/order_b
// (nullable, nl.nslots ints each): the second kernel also copies the list's
// slot order there.
void launch_pairs(bool mixed, bool energy_on, const NlistDev& nl, const PairParams& p, const ParticleDev& part,
                  const double* pos, const SysInfo* sys_dev, long long* force_fx, double* force_d,
                  long long* energy_fx, double* energy_d, unsigned* status, unsigned long long* stats,
                  int grid, hipStream_t s, bool combine = true, int* order_a = nullptr, int* order_b = nullptr);

// The second kernel alone: adds the slot-ordered pair forces into force_fx
// (mixed) or force_d and clears them (after launch_pairs with combine = false).
void launch_slot_forces(bool mixed, const NlistDev& nl, long long* force_fx, double* force_d, hipStream_t s);

// Exceptions (nonbonded.rs:995-1036), FP64 math.  has_pme selects the erf
// reciprocal correction with particle charges.
// lj_alpha > 0 adds the LJPME dispersion correction with ex.c6.
void launch_exceptions(bool mixed, const ExceptionDev& ex, const ParticleDev& part, const double* pos,
                       const SysInfo* sys_dev, int use_pbc, int has_pme, double alpha, double lj_alpha,
                       long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
                       hipStream_t s);

}  // namespace ommhip
