/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * Single-threaded FP64 C++ restatement of the reference's CPU algorithms for
 * the NonbondedForce/PME path and the water MD step around it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only
 * as the checker.  The product library never links or calls it.
 *
 * Each function cites the reference file:line it restates.  Built with
 * -O2 -ffp-contract=off so that no multiply-add is fused, matching Rust's
 * (non-contracting) f64 arithmetic operation by operation.
 *
 * Parity notes:
 *   - erfc/erf come from glibc libm, the reference uses the `libm` crate
 *     0.2.16: last-ulp values differ, parity at that level is unpinned.
 *   - The FFT is this file's own mixed-radix complex FFT with the rustfft 6.4.1
 *     conventions (forward e^{-2 pi i kn/N}, unnormalised inverse); rounding
 *     order differs from rustfft, parity at last-ulp level is unpinned.
 *   - Random streams (ChaCha12 + StandardNormal) are not restated: the
 *     integrator and barostat take their random numbers as inputs.
 */
#ifndef OMMHIP_ORACLE_H
#define OMMHIP_ORACLE_H

#include <stdint.h>
#include "../include/ommhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nonbonded.rs:647-658 (PME rule; uses the diagonal box entries). */
void orc_effective_pme_parameters(double tolerance, double cutoff, const double box[9],
                                  double* alpha, int32_t grid[3]);
/* nonbonded.rs:635-645 (LJPME rule). */
void orc_effective_ljpme_parameters(double tolerance, double cutoff, const double box[9],
                                    double* alpha, int32_t grid[3]);
/* ewald.rs:22-46. */
void orc_ewald_parameters(const double widths[3], double cutoff, double tolerance,
                          double* alpha, int32_t limits[3]);

/* NonbondedForce::evaluate (nonbonded.rs:771-1042).  The reciprocal parameters
 * are desc->pme_alpha/pme_grid when pme_alpha > 0 (the values frozen by
 * initialize), otherwise derived from `box` as the reference does before
 * initialize.  LJPME parameters: ljpme_alpha/ljpme_grid when ljpme_alpha > 0.
 * use_cells selects the cell traversal (neighbors.rs:96-181) as the reference
 * does by default; 0 forces the exhaustive i<j loop.  parts (may be NULL):
 * [0] direct pairs, [1] exceptions, [2] reciprocal mesh (incl. LJPME mesh),
 * [3] self/neutralising (+ LJPME self term), [4] dispersion correction.
 * Returns 0 on success, nonzero on a failed reference assertion. */
int orc_nb_evaluate(const omm_nb_desc* desc, double ljpme_alpha, const int32_t* ljpme_grid,
                    const double* pos, const double* box, int include_direct,
                    int include_reciprocal, int use_cells, double* forces, double* energy,
                    double parts[5]);

/* The direct-space pair set of the evaluation above: pairs (i<j) that are not
 * excluded and pass the cutoff test with the reference arithmetic
 * (nonbonded.rs:842-864).  Writes up to `capacity` pairs to out_i/out_j (may be
 * NULL), returns the total count, *hash = wrapping sum of omm_pair_key(i,j). */
int64_t orc_nb_pairs(const omm_nb_desc* desc, const double* pos, const double* box,
                     int32_t* out_i, int32_t* out_j, int64_t capacity, uint64_t* hash);

/* Candidate pairs the cell traversal visits (neighbors.rs:96-181), for the
 * superset/no-duplicate checks of neighbors.rs:227-263. */
int64_t orc_cell_candidates(const double* pos, int32_t n, double cutoff, const double* box,
                            int32_t* out_i, int32_t* out_j, int64_t capacity);

uint64_t orc_pair_key(int32_t i, int32_t j);

/* pme.rs:151-174. */
void orc_spline(double fraction, double weights[5], double derivatives[5]);
/* pme.rs:176-199; returns volume. */
double orc_reciprocal_box(const double box[9], double recip[9]);
/* Complex FFT of n points, interleaved re/im, in place; inverse unnormalised. */
void orc_fft(double* data, int32_t n, int32_t inverse);
/* B-spline moduli (pme.rs:239-259) for one axis. */
void orc_bspline_moduli(int32_t n, double* out);
/* Workspace::evaluate (pme.rs:377-425) with arbitrary amplitudes; adds to
 * forces (may be NULL) and returns the mesh energy. */
double orc_pme_mesh(const double* pos, const double* amplitudes, int32_t n, const double box[9],
                    double alpha, const int32_t grid[3], int32_t dispersion, double* forces);
/* PmeCache::charge_derivatives (pme.rs:89-138): the reciprocal potential at
 * each particle (dE_mesh/dq). */
void orc_pme_charge_derivatives(const double* pos, const double* charges, int32_t n,
                                const double box[9], double alpha, const int32_t grid[3],
                                double* derivatives);
/* EwaldCache::evaluate (ewald.rs:49-122); adds to forces, returns energy. */
double orc_ewald_reciprocal(const double* pos, const double* charges, int32_t n,
                            const double widths[3], double alpha, const int32_t limits[3],
                            double* forces);
/* dispersion_correction (nonbonded.rs:660-719) energy for this volume. */
double orc_dispersion_correction(const omm_nb_desc* desc, double volume);
/* dispersion_reciprocal_pair (nonbonded.rs:1158-1191): the positive
 * complement C6 (1 - e^-x (1 + x + x^2/2)) / r^6, x = alpha^2 r^2, and its
 * dE/dr / r. */
void orc_dispersion_reciprocal_pair(double c6, double r2, double alpha, double* energy,
                                    double* force_over_r);

/* apply_minimum_image_ortho (harmonic_bond.rs:152-159). */
void orc_minimum_image(double delta[3], const double box[9]);

/* HarmonicBondForce (harmonic_bond.rs:102-144): adds forces, returns energy. */
double orc_harmonic_bonds(int32_t nb, const int32_t* p1, const int32_t* p2, const double* length,
                          const double* k, const double* pos, const double* box, int32_t use_pbc,
                          double* forces);
/* HarmonicAngleForce (harmonic_angle.rs:102-174): adds forces, returns energy. */
double orc_harmonic_angles(int32_t na, const int32_t* p1, const int32_t* p2, const int32_t* p3,
                           const double* theta0, const double* k, const double* pos,
                           const double* box, int32_t use_pbc, double* forces);

/* LangevinMiddle host update (langevin.rs:103-142) with caller-supplied
 * standard normals (3 per massive atom, atom-major). */
void orc_lfmiddle_step(int32_t n, double* pos, double* vel, const double* forces,
                       const double* masses, double dt, double temperature, double friction,
                       const double* normals);

/* constraints.rs:35-149: SHAKE (velocities = 0) / RATTLE (velocities = 1)
 * projection of values against reference positions.  0 = converged. */
int32_t orc_constraints_solve(int32_t n, int32_t nc, const int32_t* c1, const int32_t* c2, const double* dist,
                              const double* masses, const double* reference, double* values, int32_t velocities,
                              double tolerance);
/* langevin.rs:215-277: constrained LFMiddle step (normals nullable = zero). */
int32_t orc_lfmiddle_step_constrained(int32_t n, double* pos, double* vel, const double* forces,
                                      const double* masses, double dt, double temperature, double friction,
                                      const double* normals, int32_t nc, const int32_t* c1, const int32_t* c2,
                                      const double* dist, double tolerance);
/* BOLTZMANN, langevin.rs:12. */
double orc_boltzmann(void);

/* System::molecules (system.rs:100-141) from a bond list: writes the molecule
 * index of every particle (molecules numbered by their smallest member in
 * increasing order); returns the number of molecules. */
int32_t orc_molecules(int32_t n, int32_t nbonds, const int32_t* p1, const int32_t* p2,
                      int32_t* molecule_of);
/* Context::transform_positions (context.rs:976-1029) for an isotropic scale. */
void orc_scale_positions(int32_t n, double* pos, double scale, int32_t rigid,
                         const int32_t* molecule_of, int32_t nmol);
/* MonteCarloBarostat::accept work (barostat.rs:154-178). */
/* flexible_barostat.rs:15-41: reduced form of box vectors in place; 0 = valid. */
int32_t orc_reduced_box(double* v);
/* Rigid-molecule scaling with a diagonal transform (anisotropic barostat). */
void orc_scale_positions_axes(int32_t n, double* pos, const double scale[3], const int32_t* molecule_of,
                              int32_t nmol);
double orc_barostat_work(double energy_change, double pressure_bar, double old_volume,
                         double new_volume, int32_t scaled_particles, double temperature);

#ifdef __cplusplus
}
#endif

#endif
