/*
 * ommhip.h — C ABI of the MI355X (gfx950) NonbondedForce/PME operator and the
 * water-box MD step built around it.
 *
 * The reference operator is the Rust `impl<B: Backend> Force<B> for
 * NonbondedForce` (crates/openmm-core/src/forces/nonbonded.rs:1053-1155),
 * implementing `trait Force<B>` (crates/openmm-core/src/forces/mod.rs:142-246).
 * Every entry point below names the reference item it stands in for, so a Rust
 * `impl Force<B>` (or any other FFI host) can forward to it one call per method.
 *
 * Conventions (same as the reference):
 *   - units nm, ps, kJ/mol, elementary charge, Da;
 *   - positions/forces are N x 3 row-major doubles (x0 y0 z0 x1 ...);
 *   - box vectors are 9 doubles, rows a, b, c, in reduced triclinic form
 *     (crates/openmm-core/src/system.rs:16-39); NULL means non-periodic;
 *   - forces are returned in a caller-owned buffer that is OVERWRITTEN (the
 *     reference returns a new owned tensor that the Context sums).
 *
 * Errors: every function returns an omm_status; nothing unwinds across the
 * ABI.  omm_last_error() returns a thread-local message for the last failure.
 * A handle is used by one host thread at a time (calls on one handle are
 * serialised by an internal mutex, matching the reference's Mutex-guarded
 * caches, pme.rs:10, neighbors.rs:8).
 */
#ifndef OMMHIP_H
#define OMMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMMHIP_ABI_VERSION 3

typedef enum omm_status {
    OMM_OK = 0,
    OMM_ERR_INVALID = 1,     /* bad argument / descriptor (reference: assert!) */
    OMM_ERR_CUTOFF_BOX = 2,  /* cutoff > half the box (nonbonded.rs:1080-1085, 806-812) */
    OMM_ERR_NONFINITE = 3,   /* non-finite position or energy (neighbors.rs:108-111) */
    OMM_ERR_OVERFLOW = 4,    /* fixed-point or neighbour-list capacity overflow */
    OMM_ERR_DEVICE = 5,      /* HIP runtime error or no usable gfx950 device */
    OMM_ERR_STATE = 6        /* call out of order (e.g. compute before initialize) */
} omm_status;

/* NonbondedMethod, nonbonded.rs:20-33. */
typedef enum omm_nb_method {
    OMM_NB_NO_CUTOFF = 0,
    OMM_NB_CUTOFF_NON_PERIODIC = 1,
    OMM_NB_CUTOFF_PERIODIC = 2,
    OMM_NB_EWALD = 3,
    OMM_NB_PME = 4,
    OMM_NB_LJPME = 5
} omm_nb_method;

/* Arithmetic of the device path.
 *   MIXED : FP32 pair/spline/FFT math, int64 fixed-point force/grid/energy
 *           accumulation, FP64 positions, integrator and final reductions.
 *   DOUBLE: the plain FP64 form — same algorithms in FP64 throughout, FP64
 *           atomics, one launch sequence per system and no graph. */
typedef enum omm_precision {
    OMM_PRECISION_MIXED = 0,
    OMM_PRECISION_DOUBLE = 1
} omm_precision;

/*
 * Description of a NonbondedForce: the builder state of the reference struct
 * (nonbonded.rs:75-108) as plain arrays.  Setter equivalents:
 * add_particle (165), add_exception (197), set_nonbonded_method (502),
 * set_cutoff_distance (520), set_reaction_field_dielectric (533),
 * set_ewald_error_tolerance (545), set_exceptions_use_periodic_boundary_conditions
 * (557), set_use_dispersion_correction (570), set_use_switching_function (576),
 * set_switching_distance (583), set_pme_parameters (595), set_force_group
 * (forces/mod.rs:58), set_reciprocal_space_force_group (115),
 * set_include_direct_space (125).  The arrays are copied at create time.
 */
typedef struct omm_nb_desc {
    int32_t num_particles;
    const double* charge;    /* [num_particles] e */
    const double* sigma;     /* [num_particles] nm */
    const double* epsilon;   /* [num_particles] kJ/mol */
    int32_t num_exceptions;
    const int32_t* exception_particle1;  /* [num_exceptions] */
    const int32_t* exception_particle2;
    const double* exception_charge_prod;
    const double* exception_sigma;
    const double* exception_epsilon;
    int32_t method;                      /* omm_nb_method */
    double cutoff;                       /* nm, default 1.0 */
    double reaction_field_dielectric;    /* default 78.3 */
    double ewald_error_tolerance;        /* default 5e-4 */
    double pme_alpha;                    /* 0 = derive from tolerance (nonbonded.rs:647-658) */
    int32_t pme_grid[3];
    int32_t use_dispersion_correction;   /* default 1 */
    int32_t exceptions_use_pbc;          /* default 0 */
    int32_t use_switching_function;      /* default 0 */
    double switching_distance;           /* default -1 */
    int32_t force_group;                 /* 0..31, default 0 */
    int32_t reciprocal_force_group;      /* -1..31, default -1 (= force_group) */
    int32_t include_direct_space;        /* default 1 */
    /* Global parameters and parameter offsets (nonbonded.rs:247-430).  The
     * parameters used are base + value(parameter) * scale summed over the
     * offsets of each particle / exception, a later offset of the same
     * (parameter, target) replacing an earlier one (nonbonded.rs:395-430).
     * All counts default to 0. */
    int32_t num_global_parameters;
    const char* const* global_parameter_names;   /* non-empty, distinct */
    const double* global_parameter_defaults;     /* finite */
    int32_t num_particle_offsets;
    const int32_t* particle_offset_parameter;    /* index into the global parameters */
    const int32_t* particle_offset_index;        /* particle */
    const double* particle_offset_scales;        /* [n][3]: charge, sigma, epsilon */
    int32_t num_exception_offsets;
    const int32_t* exception_offset_parameter;
    const int32_t* exception_offset_index;       /* exception */
    const double* exception_offset_scales;       /* [n][3]: charge product, sigma, epsilon */
    /* LJPME dispersion mesh (nonbonded.rs:615-645): 0 = derive alpha and the
     * grid from the tolerance (half the Coulomb rule's grid). */
    double ljpme_alpha;
    int32_t ljpme_grid[3];
} omm_nb_desc;

/* Fill *desc with the reference defaults of NonbondedForce::new (nonbonded.rs:129-162);
 * array pointers are set to NULL and counts to 0. */
void omm_nb_desc_defaults(omm_nb_desc* desc);

typedef struct omm_nb omm_nb;

/* Thread-local text of the last error on this thread ("" if none). */
const char* omm_last_error(void);

/* OMMHIP_ABI_VERSION of the loaded library. */
int32_t omm_abi_version(void);

/* Number of visible gfx950 devices (0 when none; never an error on a host
 * without a GPU). */
int32_t omm_device_count(void);

/* Validates and copies the description; allocates no device memory yet.
 * Mirrors NonbondedForce::new + the builders above; the parameter checks are
 * validate_particle_parameters (nonbonded.rs:1252-1261) and the setter asserts. */
omm_status omm_nb_create(const omm_nb_desc* desc, int32_t precision, omm_nb** out);

/* Force::initialize (nonbonded.rs:1055-1107): checks exception indices and
 * duplicates, switching distance, cutoff <= box/2, then freezes the PME alpha
 * and grid from this box for the lifetime of the handle.  Allocates the device
 * buffers.  box may be NULL for non-periodic methods. */
omm_status omm_nb_initialize(omm_nb* nb, int32_t num_particles, const double* box);

/* The frozen (alpha, grid) — NonbondedForce::pme_parameters after initialize
 * (initialized_reciprocal, nonbonded.rs:1086-1102). */
omm_status omm_nb_get_pme_parameters(const omm_nb* nb, double* alpha, int32_t grid[3]);
/* NonbondedForce LJPME parameters frozen by initialize (nonbonded.rs:1098-1100,
 * 635-645): alpha and grid of the dispersion mesh; zeros for other methods. */
omm_status omm_nb_get_ljpme_parameters(const omm_nb* nb, double* alpha, int32_t grid[3]);

/* Force::calc_forces_and_energy (nonbonded.rs:1125-1131): direct space if
 * include_direct_space, reciprocal always.  pos: host N x 3; box: host 9 or
 * NULL; forces_out: host N x 3 (overwritten, may be NULL); energy_out: kJ/mol.
 * With energy_out == NULL only forces are evaluated, on the MD step's path
 * (MIXED: FP32 pair and mesh arithmetic); with an energy requested, MIXED
 * evaluates each pair energy and the whole mesh in FP64 as well. */
omm_status omm_nb_calc_forces_and_energy(omm_nb* nb, const double* pos, const double* box,
                                         double* forces_out, double* energy_out);

/* Force::calc_forces_and_energy_groups (nonbonded.rs:1133-1150).  *evaluated
 * is set to 0 (and outputs untouched) when neither the direct nor the
 * reciprocal group is selected — the reference's `None`. */
omm_status omm_nb_calc_forces_and_energy_groups(omm_nb* nb, const double* pos, const double* box,
                                                uint32_t groups, double* forces_out,
                                                double* energy_out, int32_t* evaluated);

/* The resident form of calc_forces_and_energy_groups (no counterpart in the
 * reference, which takes host tensors; for callers whose positions already
 * live on the device, e.g. a device-side integrator).  d_pos: device N x 3,
 * read after the work already queued on `stream` (a hipStream_t; NULL: the
 * null stream); d_forces: device N x 3 (overwritten, may be NULL); box and
 * energy_out on the host as above; groups as above (0xFFFFFFFF: all).  The
 * call returns once the outputs are complete.  Positions are not copied to
 * the host except for NoCutoff / CutoffNonPeriodic, whose virtual box is
 * chosen from them. */
omm_status omm_nb_calc_forces_and_energy_device(omm_nb* nb, const double* d_pos, const double* box,
                                                uint32_t groups, double* d_forces, double* energy_out,
                                                int32_t* evaluated, void* stream);

/* Energy split of the last evaluation (same call as above, with
 * include_direct = include_reciprocal = 1):
 *   parts[0] direct-space pairs (LJ + screened/plain Coulomb),
 *   parts[1] exceptions incl. the erf reciprocal correction (nonbonded.rs:995-1036),
 *   parts[2] reciprocal-space mesh energy (pme.rs:339-372),
 *   parts[3] self + neutralising energy (nonbonded.rs:965-973),
 *   parts[4] dispersion correction (nonbonded.rs:991-993).
 * Their sum is the energy of omm_nb_calc_forces_and_energy. */
omm_status omm_nb_energy_parts(omm_nb* nb, const double* pos, const double* box, double parts[5]);

/* Diagnostics for neighbour/exclusion-set parity: the number of direct-space
 * pairs (i<j, not excluded, r^2 < cutoff^2 with the reference's FP64
 * minimum-image arithmetic, nonbonded.rs:842-864) that the device evaluated in
 * the last call, and an order-independent 64-bit hash of the set: the wrapping
 * sum over pairs of omm_pair_key(i, j). */
omm_status omm_nb_pair_stats(omm_nb* nb, const double* pos, const double* box,
                             int64_t* num_pairs, uint64_t* pair_hash);

/* Mixing function of the pair hash (splitmix64 finaliser of (i << 32) | j). */
uint64_t omm_pair_key(int32_t i, int32_t j);

/* Frees everything owned by the handle. NULL is accepted. */
/* Force::set_global_parameter (nonbonded.rs:1113-1123): *recognised = 1 when
 * the force has the parameter (then its value is set and the effective
 * parameters follow), 0 otherwise. */
omm_status omm_nb_set_global_parameter(omm_nb* nb, const char* name, double value, int32_t* recognised);

/* Current value of a global parameter (*recognised as above). */
omm_status omm_nb_get_global_parameter(omm_nb* nb, const char* name, double* value, int32_t* recognised);

/* NonbondedForce::update_parameters_from (nonbonded.rs:721-767): new particle
 * and exception parameters, offsets and global parameter defaults from desc,
 * under the reference's rules (same counts, same exception pairs, same
 * offset targets, excluded exceptions stay excluded).  Current global
 * parameter values are kept. */
omm_status omm_nb_update_parameters(omm_nb* nb, const omm_nb_desc* desc);

omm_status omm_nb_destroy(omm_nb* nb);

/* ------------------------------------------------------------------------ */
/* The charge mesh as a second consumer of the device PME: the crate-internal
 * PmeCache methods that ConstantPotentialForce calls (pme.rs:63-138; called at
 * constant_potential.rs:642-644 for its energy and forces, and by its charge
 * solver, constant_potential/solver.rs:24-31 and 219-226, for dE/dq).  Each
 * call brings positions (n x 3), charges (n), box vectors (9, reduced form),
 * alpha and the grid (each >= 5), as the reference's calls do; a workspace is
 * set up for each (n, grid) and the two most recent are kept.  The mesh is
 * plain FP64 (FP64 grid, transforms and sums): its users solve linear
 * systems with it.  Uncharged particles do not spread (pme.rs:318). */
typedef struct omm_pme omm_pme;

omm_status omm_pme_create(omm_pme** out);

/* PmeCache::evaluate_charges (pme.rs:63-86): the reciprocal mesh energy of
 * these charges (*energy) and its forces, ADDED to forces (n x 3, nullable)
 * as the reference adds into its force buffer. */
omm_status omm_pme_evaluate_charges(omm_pme* pme, int32_t num_particles, const double* pos, const double* charges,
                                    const double* box, double alpha, const int32_t grid[3], double* forces,
                                    double* energy);

/* PmeCache::charge_derivatives (pme.rs:88-138): for each of the count
 * particle indices, the mesh potential at that particle's stencil (the
 * derivative of the mesh energy with respect to its charge) is ADDED to
 * derivatives[k]. */
omm_status omm_pme_charge_derivatives(omm_pme* pme, int32_t num_particles, const double* pos, const double* charges,
                                      const double* box, double alpha, const int32_t grid[3],
                                      const int32_t* indices, int32_t count, double* derivatives);

omm_status omm_pme_destroy(omm_pme* pme);

/* No counterpart in the reference: omm_pme_charge_derivatives for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  It calls no hipMalloc, hipFree or allocating library routine.
 * pos (n x 3), charges (n), indices (count) and derivatives (count, ADDED to)
 * are the caller's DEVICE buffers; box (9), alpha and grid are host values.
 * Every kernel is launched on `stream` (a hipStream_t; NULL: the default
 * stream) and the call returns without waiting.  The only other device
 * memory is a static FP64 complex grid of 64^3 points (4 MiB) per device,
 * held until the process ends: grids above 64 points on an axis are rejected
 * with OMM_ERR_INVALID.  Every call on a device shares that grid (each call
 * clears it first), so calls that overlap on one device, from other streams
 * or threads, race.  An index outside [0, n) sets its derivative to NaN. */
omm_status omm_pme_charge_derivatives_noalloc(int32_t num_particles, const double* pos, const double* charges,
                                              const double* box, double alpha, const int32_t grid[3],
                                              const int32_t* indices, int32_t count, double* derivatives,
                                              void* stream);

/* ------------------------------------------------------------------------ */
/* MD context: the step around the operator for the flexible-water workloads.
 * Stands in for Context::step (context.rs:795-928) with a LangevinMiddle
 * integrator (integrators/langevin.rs:103-142), HarmonicBondForce
 * (harmonic_bond.rs:93-148), HarmonicAngleForce (harmonic_angle.rs:99-186) and
 * the NonbondedForce above, all resident on the device.  Random numbers come
 * from a counter-based Philox4x32-10 stream keyed by (seed, step, atom), so the
 * trajectory is reproducible and resumable but not the reference's ChaCha12
 * stream (the reference notes random paths differ across implementations). */

typedef struct omm_md_desc {
    int32_t num_particles;
    const double* masses;              /* [N] Da; 0 = fixed particle */
    const omm_nb_desc* nonbonded;
    int32_t num_bonds;
    const int32_t* bond_particle1;
    const int32_t* bond_particle2;
    const double* bond_length;         /* nm */
    const double* bond_k;              /* kJ/mol/nm^2 */
    int32_t num_angles;
    const int32_t* angle_particle1;
    const int32_t* angle_particle2;    /* vertex */
    const int32_t* angle_particle3;
    const double* angle_theta0;        /* rad */
    const double* angle_k;             /* kJ/mol/rad^2 */
    double box[9];
    double temperature;                /* K */
    double friction;                   /* 1/ps */
    double step_size;                  /* ps */
    uint64_t seed;
    /* Distance constraints (System::add_constraint; constraints.rs): SHAKE
     * positions / RATTLE velocities inside the LangevinMiddle step
     * (langevin.rs:215-277).  0 constraints = unconstrained (default). */
    int32_t num_constraints;
    const int32_t* constraint_particle1;
    const int32_t* constraint_particle2;
    const double* constraint_distance;  /* nm */
    double constraint_tolerance;        /* 0 = the integrator default 1e-5 (langevin.rs:63) */
} omm_md_desc;

typedef struct omm_ctx omm_ctx;

/* Context::new (context.rs:83-161) for the description above: initialises the
 * nonbonded operator (freezing alpha/grid) and uploads the topology. */
omm_status omm_ctx_create(const omm_md_desc* desc, int32_t precision, omm_ctx** out);

/* Context::set_positions / set_velocities (host N x 3). */
omm_status omm_ctx_set_positions(omm_ctx* ctx, const double* pos);
omm_status omm_ctx_set_velocities(omm_ctx* ctx, const double* vel);

/* Context::apply_constraints / apply_velocity_constraints (context.rs:767-810):
 * project the positions onto the constraints (against themselves), or the
 * velocities (against the positions), with the given tolerance. */
omm_status omm_ctx_apply_constraints(omm_ctx* ctx, double tolerance);
omm_status omm_ctx_apply_velocity_constraints(omm_ctx* ctx, double tolerance);

/* Context::set_velocities_to_temperature (context.rs:182-223): Maxwell-Boltzmann
 * velocities from `seed`, minus F/m * dt/2 (the LangevinMiddle velocity time
 * offset, langevin.rs:204-206). */
omm_status omm_ctx_set_velocities_to_temperature(omm_ctx* ctx, double temperature, uint64_t seed);

/* LangevinMiddleIntegrator::set_temperature / set_friction / set_step_size
 * (integrators/langevin.rs:82-101) on the context's integrator; same checks
 * (temperature and friction nonnegative and finite, step size positive). */
omm_status omm_ctx_set_integrator(omm_ctx* ctx, double temperature, double friction, double step_size);

/* Context::set_parameter (context.rs:251-261) for the NonbondedForce's global
 * parameters: *recognised = 1 when the force has the parameter. */
omm_status omm_ctx_set_parameter(omm_ctx* ctx, const char* name, double value, int32_t* recognised);

/* Context::step (context.rs:795): nsteps LangevinMiddle steps.  The mixed
 * form replays a captured HIP graph of 10 steps; each step's integrator
 * flags the neighbour list when an atom moved more than half the list skin,
 * and the next step's build kernels rebuild it on the device. */
omm_status omm_ctx_step(omm_ctx* ctx, int32_t nsteps);

/* Steps per captured HIP graph of the mixed form (default 10; 0 launches
 * every step eagerly, with the same kernels and results).  Launch policy, no
 * counterpart in the reference. */
omm_status omm_ctx_set_graph_steps(omm_ctx* ctx, int32_t steps);

/* MonteCarloBarostat (forces/barostat.rs:59-218) added to the context:
 * isotropic volume moves every `frequency` steps at `pressure` bar and
 * `temperature` K, molecules (bond-connected components, system.rs:100-141)
 * scaled as rigid bodies; at frequency 0 it makes no attempts but stays set
 * for the pressure queries.  The attempt runs at the
 * start of the step that completes a period (context/program.rs:5-45), with
 * energies from the FP64 energy path.  The random stream (Philox keyed by
 * seed) differs from the reference's ChaCha12, as the reference allows for
 * seeded streams across implementations.  alpha and the PME grid stay as
 * initialised (nonbonded.rs:1055-1107). */
omm_status omm_ctx_set_barostat(omm_ctx* ctx, double pressure, double temperature, int32_t frequency,
                                uint64_t seed);

/* Barostat bookkeeping of one system (no counterpart in the reference, which
 * keeps it private): the adaptive step and counters (barostat.rs:180-195)
 * and the record of the last attempt. */
typedef struct omm_barostat_info {
    double volume_scale;
    int64_t attempted, accepted;              /* since the last step-size change */
    int64_t total_attempted, total_accepted;
    double volume_old, volume_new;            /* nm^3 */
    double energy_old, energy_new, work;      /* kJ/mol */
    int32_t accepted_last, skipped;
    int32_t axis;                             /* anisotropic / membrane: axis of the last attempt */
    double scale[3];                          /* box scale factors of the last proposal */
    double area_change;                       /* membrane: xy area change of the last proposal (nm^2) */
    double log_shape;                         /* flexible: 2 ln s_x + ln s_y of the last proposal */
    double box_new[9];                        /* the last proposal's box vectors */
} omm_barostat_info;

omm_status omm_ctx_barostat_info(omm_ctx* ctx, omm_barostat_info* info);

/* Context::compute_current_pressure (context/pressure.rs:14-42) for the
 * MonteCarloBarostat set on the context (OMM_ERR_STATE without one), in bar:
 * the FP64 potential energies at the box scaled by 1.001 and 0.999 with the
 * molecules moved rigidly (context.rs:942-1029), and the kinetic energy of
 * the molecular momenta (pressure.rs:107-131).  Positions, velocities, box
 * and step count are unchanged. */
omm_status omm_ctx_compute_current_pressure(omm_ctx* ctx, double* pressure);

/* Context::compute_current_anisotropic_pressure and
 * compute_current_membrane_pressure (pressure.rs:44-105): the pressure along
 * each Cartesian axis (bar), for the barostat of that kind. */
omm_status omm_ctx_compute_current_anisotropic_pressure(omm_ctx* ctx, double pressure[3]);
omm_status omm_ctx_compute_current_membrane_pressure(omm_ctx* ctx, double pressure[3]);

/* Context::compute_current_flexible_pressure (pressure.rs:133-170) and
 * compute_stress_tensor (pressure.rs:172-209) for the flexible barostat:
 * components xx yy zz xy xz yz (bar), each from trial boxes deformed by
 * +-0.1 % and put in reduced form. */
omm_status omm_ctx_compute_current_flexible_pressure(omm_ctx* ctx, double pressure[6]);
omm_status omm_ctx_compute_stress_tensor(omm_ctx* ctx, int32_t include_kinetic, double stress[6]);

/* MonteCarloAnisotropicBarostat (forces/anisotropic_barostat.rs:42-243):
 * pressures per axis (bar), scale_axes bit d enables axis d; each attempt
 * scales one enabled axis by V'/V with its own adaptive step.  Replaces any
 * barostat set before; frequency 0 stops the attempts. */
omm_status omm_ctx_set_anisotropic_barostat(omm_ctx* ctx, const double pressure[3], double temperature,
                                            int32_t scale_axes, int32_t frequency, uint64_t seed);

/* MonteCarloFlexibleBarostat (forces/flexible_barostat.rs:132-238): every
 * lower-triangle box element moves by up to an adaptive length step; the
 * trial box is reduced and molecules scale by the diagonal ratios.
 * (volume_scale in omm_barostat_info then reports the length step.) */
omm_status omm_ctx_set_flexible_barostat(omm_ctx* ctx, double pressure, double temperature, int32_t frequency,
                                         uint64_t seed);

/* MonteCarloMembraneBarostat (forces/membrane_barostat.rs:61-292): pressure
 * (bar), surface tension (bar nm), xy_mode 0 isotropic / 1 anisotropic,
 * z_mode 0 free / 1 fixed / 2 constant volume. */
omm_status omm_ctx_set_membrane_barostat(omm_ctx* ctx, double pressure, double surface_tension, double temperature,
                                         int32_t xy_mode, int32_t z_mode, int32_t frequency, uint64_t seed);

/* Current box vectors (row-major a, b, c); they change under a barostat
 * (Context::get_state box vectors, context.rs:930-1105). */
omm_status omm_ctx_get_box(omm_ctx* ctx, double box[9]);

/* One barostat attempt now, outside the step sequence (test hook; no
 * counterpart in the reference). */
omm_status omm_ctx_barostat_attempt(omm_ctx* ctx);

/* Shape of the current neighbour list (measurement hook; no counterpart in the
 * reference): number of 8 x 64 tiles (8-atom cluster x 64 list entries) the
 * pair kernel evaluates and the total number of cluster-list entries.
 * Pair-slot efficiency = pairs / (tiles * 512). */
omm_status omm_ctx_list_stats(omm_ctx* ctx, int64_t* tiles, int64_t* entries);

/* Storage test hooks (no counterpart in the reference, whose cell list
 * grows on the host): omm_ctx_set_capacity reallocates the neighbour-list
 * storage to list_cap entries per cluster (rounded up to whole 64-entry tiles)
 * and rebuilds the lists from the positions they were built from, and the PME
 * spread bins to bin_cap atoms per tile; 0 keeps a capacity.  A capacity
 * below the need truncates without writing out of bounds and raises the
 * overflow status; the next omm_ctx_step then restores the state it started
 * from, grows the storage and replays, ending on the trajectory of a run that
 * never overflowed.  omm_ctx_capacity reports the capacities and how many
 * runs of steps were replayed so. */
omm_status omm_ctx_set_capacity(omm_ctx* ctx, int32_t list_cap, int32_t bin_cap);
omm_status omm_ctx_capacity(omm_ctx* ctx, int32_t* list_cap, int32_t* bin_cap, int64_t* overflow_replays);

/* Launch shapes picked by the autotuner when the context is first prepared
 * (no counterpart in the reference): the spread's decomposition (the tile
 * edge in grid points for tiles of the grid, or 0 for chunks of sorted atoms
 * summed in LDS boxes, mixed precision only), the FFT decomposition (1: the z and y transforms of each x plane in one
 * workgroup, three launches per transform; 0: five line passes) with the
 * lines per workgroup of the line passes, and the pair kernel's grid (the
 * calibrated task count times pair_scale: 1, 2, or -2 for half), each the
 * fastest of its candidates on the context's own positions, with that
 * candidate's time.  None of them changes a result bit.  from_cache = 1 when
 * the shapes came from the tuning cache (omm_set_tuning_cache) instead of
 * being timed; the times are then those of the run that stored them. */
typedef struct omm_launch_tuning {
    int32_t spread_tile_edge;
    int32_t fft_lines_per_workgroup;
    int32_t pair_grid;
    double spread_ms, fft_ms, pair_ms;
    int32_t fft_plane;
    int32_t pair_scale;
    int32_t from_cache;
} omm_launch_tuning;

omm_status omm_ctx_launch_tuning(omm_ctx* ctx, omm_launch_tuning* out);

/* Sets the launch shapes instead of the autotuner's (no counterpart in the
 * reference): spread_tile_edge 6 or 8, or 0 for the atom-owned spread (mixed
 * precision only), fft_plane 0 or 1 (1 only where every
 * system's x plane fits in LDS), fft_lines_per_workgroup 16, 8, 4 or 2,
 * pair_scale 1, 2 or -2; the other fields are ignored.  Results do not
 * change; OMM_ERR_INVALID for other values. */
omm_status omm_ctx_set_launch_tuning(omm_ctx* ctx, const omm_launch_tuning* tuning);

/* The process-wide tuning cache (no counterpart in the reference): a text
 * file keyed by device (architecture and compute units) and workload
 * (precision, method, cutoff, system sizes and grids).  A context or batch
 * whose key is in the file applies the stored shapes without timing; every
 * timed tuning appends its result.  NULL or "" turns the cache off.  Without
 * this call the path comes from the environment variable
 * OMMHIP_TUNING_CACHE, if set. */
omm_status omm_set_tuning_cache(const char* path);

/* Context::get_state (context.rs:930-1105) with positions, velocities, forces
 * and potential energy; any output pointer may be NULL. */
omm_status omm_ctx_get_state(omm_ctx* ctx, double* pos, double* vel, double* forces,
                             double* potential_energy, int64_t* step_count);

/* Context::set_state (context.rs:274-293): step count, box vectors (9,
 * NULL: kept), positions and velocities (N x 3, NULL: kept).  Global
 * parameters are restored with omm_ctx_set_parameter, as the reference's
 * set_state loops over set_parameter.  The step count keys the integrator's
 * noise, so a restored context draws what the saved one would have drawn.
 * list_positions (N x 3, NULL: pos) are the positions the saved context last
 * built its neighbour list from (omm_ctx_get_list_positions); the list is
 * rebuilt from them, which reproduces the saved list, so the restored context
 * continues the saved trajectory bit for bit.  Without them the list is built
 * from pos: the same physics, with FP32 pair sums taken around other block
 * centres.  Barostat bookkeeping stays with the barostat, as in the reference. */
omm_status omm_ctx_set_state(omm_ctx* ctx, int64_t step_count, const double* box, const double* pos,
                             const double* vel, const double* list_positions);

/* Positions the current neighbour list was built from (N x 3): the list part
 * of a checkpoint, for omm_ctx_set_state (no counterpart in the reference,
 * whose cell list is rebuilt on every evaluation). */
omm_status omm_ctx_get_list_positions(omm_ctx* ctx, double* list_positions);

/* Kinetic energy (context.rs:1107-1163), kJ/mol, of the current velocities. */
omm_status omm_ctx_kinetic_energy(omm_ctx* ctx, double* kinetic_energy);

/* Measurement hook (no counterpart in the reference, which publishes no
 * per-phase timing): average device time in ms of each phase of the step over
 * `reps` launches, timed with HIP events on the context's stream:
 *   [0] neighbour-list rebuild, [1] direct-space pairs, [2] exceptions,
 *   [3] PME spread, [4] PME FFT passes + convolution, [5] PME gather,
 *   [6] bonds + angles, [7] LangevinMiddle update.
 * *num_pairs (nullable) receives the in-cutoff pair count of the current
 * positions.  Positions and velocities are left unchanged. */
omm_status omm_ctx_profile(omm_ctx* ctx, int32_t reps, double phase_ms[8], int64_t* num_pairs);

omm_status omm_ctx_destroy(omm_ctx* ctx);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic flexible 3-site water box (host only, no device use).
 * num_molecules waters: O on a jittered cubic lattice (m = ceil(M^(1/3)) sites
 * per axis, spacing L/m, jitter U(+-0.02 nm)), H at r(OH)=0.09572 nm and
 * HOH = 104.52 deg, uniformly random orientation (Shoemake quaternion).
 * Random stream: splitmix64-seeded xoshiro256**.  Atom order O, H1, H2 per
 * molecule.  positions: [3*num_molecules*3] nm.  Charges/LJ/masses are the
 * fixed water parameters (O: -0.834 e, sigma 0.315061 nm, eps 0.636386 kJ/mol,
 * 15.999 Da; H: +0.417 e, sigma 1.0, eps 0, 1.008 Da). */
omm_status omm_water_box(int32_t num_molecules, double box_edge, uint64_t seed, double* positions);

/* ---------------------------------------------------------------------------
 * Batched replicas.  The reference runs one Context per independent system;
 * a batch holds many such systems (each with its own particles, topology,
 * box, PME grid and seed) and advances all of them with one set of launches
 * per phase inside one captured graph.  Replica r of a batch follows the same
 * trajectory bit for bit as an omm_ctx created from the same descriptor
 * (forces are fixed-point sums; the RNG is keyed by the replica's seed).
 * The replicas must share the nonbonded method, cutoff, tolerance and
 * options (so alpha is shared) and the integrator settings.
 * ------------------------------------------------------------------------ */

typedef struct omm_batch omm_batch;

/* One Context::new (context.rs:83-161) per descriptor of replicas[0..count). */
omm_status omm_batch_create(const omm_md_desc* replicas, int32_t count, int32_t precision, omm_batch** out);
omm_status omm_batch_size(omm_batch* b, int32_t* count);

/* Context::set_positions / set_velocities of replica r (host N_r x 3). */
omm_status omm_batch_set_positions(omm_batch* b, int32_t replica, const double* pos);
omm_status omm_batch_set_velocities(omm_batch* b, int32_t replica, const double* vel);

/* Context::set_velocities_to_temperature of every replica; replica r draws
 * from seed + r. */
omm_status omm_batch_set_velocities_to_temperature(omm_batch* b, double temperature, uint64_t seed);

/* LangevinMiddleIntegrator settings shared by every replica (as
 * omm_ctx_set_integrator). */
omm_status omm_batch_set_integrator(omm_batch* b, double temperature, double friction, double step_size);

/* Context::step of every replica (nsteps each). */
omm_status omm_batch_step(omm_batch* b, int32_t nsteps);

/* As omm_ctx_set_graph_steps, for the whole batch. */
omm_status omm_batch_set_graph_steps(omm_batch* b, int32_t steps);

/* As omm_ctx_set_barostat for every replica; replica r draws from seed + r. */
omm_status omm_batch_set_barostat(omm_batch* b, double pressure, double temperature, int32_t frequency,
                                  uint64_t seed);
omm_status omm_batch_barostat_info(omm_batch* b, int32_t replica, omm_barostat_info* info);
omm_status omm_batch_set_anisotropic_barostat(omm_batch* b, const double pressure[3], double temperature,
                                              int32_t scale_axes, int32_t frequency, uint64_t seed);
omm_status omm_batch_set_flexible_barostat(omm_batch* b, double pressure, double temperature, int32_t frequency,
                                           uint64_t seed);
omm_status omm_batch_set_membrane_barostat(omm_batch* b, double pressure, double surface_tension,
                                           double temperature, int32_t xy_mode, int32_t z_mode, int32_t frequency,
                                           uint64_t seed);
omm_status omm_batch_barostat_attempt(omm_batch* b);

/* The pressure queries of every replica at once (the same trial evaluations
 * serve all replicas): count values, or 3 / 6 per replica, replica-major. */
omm_status omm_batch_compute_current_pressure(omm_batch* b, double* pressure);
omm_status omm_batch_compute_current_anisotropic_pressure(omm_batch* b, double* pressure);
omm_status omm_batch_compute_current_membrane_pressure(omm_batch* b, double* pressure);
omm_status omm_batch_compute_current_flexible_pressure(omm_batch* b, double* pressure);
omm_status omm_batch_compute_stress_tensor(omm_batch* b, int32_t include_kinetic, double* stress);

/* Context::get_state of replica r; any output pointer may be NULL. */
omm_status omm_batch_get_state(omm_batch* b, int32_t replica, double* pos, double* vel, double* forces,
                               double* potential_energy, int64_t* step_count);

/* Context::set_state of replica r (as omm_ctx_set_state).  The step count is
 * shared by the replicas of a batch (it keys every replica's noise stream):
 * the last value set holds for all of them. */
omm_status omm_batch_set_state(omm_batch* b, int32_t replica, int64_t step_count, const double* box,
                               const double* pos, const double* vel, const double* list_positions);
omm_status omm_batch_get_list_positions(omm_batch* b, int32_t replica, double* list_positions);

/* Potential energy of every replica (count values), one device evaluation. */
omm_status omm_batch_energies(omm_batch* b, double* potential_energies);

/* Kinetic energy of replica r, kJ/mol. */
omm_status omm_batch_kinetic_energy(omm_batch* b, int32_t replica, double* kinetic_energy);

/* Current box vectors of replica r (row-major a, b, c). */
omm_status omm_batch_get_box(omm_batch* b, int32_t replica, double box[9]);

/* Measurement hooks (no counterpart in the reference): list shape and the
 * per-phase device times of omm_ctx_profile, over the whole batch. */
omm_status omm_batch_list_stats(omm_batch* b, int64_t* tiles, int64_t* entries);
omm_status omm_batch_profile(omm_batch* b, int32_t reps, double phase_ms[8], int64_t* num_pairs);
omm_status omm_batch_launch_tuning(omm_batch* b, omm_launch_tuning* out);
omm_status omm_batch_set_launch_tuning(omm_batch* b, const omm_launch_tuning* tuning);

omm_status omm_batch_destroy(omm_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* OMMHIP_H */
