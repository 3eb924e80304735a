// Smooth PME reciprocal space (order-5 B-splines), pme.rs:151-425, for a
// batch of systems with their own grids: one launch per phase serves every
// system (spread and gather by 64-slot chunks of the sorted order, each FFT
// pass by groups of lines), the per-system layout coming from SysInfo.
#pragma once

#include "pairs.h"

namespace ommhip {

struct PmeDev {
    int nsys = 0;
    long long qpoints = 0, cpoints = 0;  // real / half-spectrum points over all systems
    DevBuf<long long> qgrid_fx;  // mixed: fixed-point charge grids
    DevBuf<double> qgrid_d;      // plain: FP64 charge grids
    DevBuf<float> cgrid_f;       // complex half-spectra [nx][ny][nzh] (interleaved re, im)
    DevBuf<double> cgrid_d;
    DevBuf<float> phi_f;         // real-space potentials [nx][ny][nz]
    DevBuf<double> phi_d;
    DevBuf<float> roots_f;       // per system and axis n roots e^{-2 pi i m/n}, interleaved
    DevBuf<double> roots_d;
    DevBuf<float> moduli_f;      // per system B-spline moduli, x then y then z
    DevBuf<double> moduli_d;
    // Workgroup prefix tables of the five FFT passes, per precision:
    // fft_off[prec][pass] has nsys + 1 entries (prec 0: FP32, 1: FP64).
    DevBuf<int> fft_off[2][5];
    int fft_blocks[2][5] = {};
    int fft_cap[2] = {};  // lines per FFT workgroup (at most), per precision
    // Plane decomposition: the z and y transforms of each x plane in one
    // workgroup (forward: z r2c then y; inverse: y then z c2r), so a
    // transform is three launches instead of five.  Available per precision
    // when every system's plane fits in LDS; results are the same bits as
    // the five-pass form.
    DevBuf<int> plane_off;  // nsys + 1 prefix of x planes
    int plane_blocks = 0;
    size_t plane_shmem[2] = {};  // dynamic LDS per plane workgroup, 0: unavailable
    bool plane[2] = {};          // use the plane decomposition
    int spread_edge = 0;  // spread tile edge
    // Atom-owned spread (chunks of sorted atoms summed in LDS boxes) instead
    // of the tile-owned one; both write the same bits.
    bool spread_atoms = false;
    size_t fft_shmem[2][3] = {};  // dynamic LDS per axis
    // Spread: per-atom stencils (base point, weights) and per-tile atom bins.
    int natoms = 0, ntiles = 0, bin_cap = 0;
    double max_charge = 0.0;        // largest |q| (selects the spread's int32 rounds)
    const int* atom_sys = nullptr;  // system of each atom (owned by the caller)
    DevBuf<int> tile_off;           // nsys + 1 prefix of tiles per system
    DevBuf<int> bin_count, bins;    // [ntiles], [ntiles][bin_cap]
    DevBuf<int> st_idx;             // [natoms][4] stencil base point
    DevBuf<float> st_wf;            // [natoms][16] weights, FP32 mesh
    DevBuf<double> st_wd;           // [natoms][16] weights, FP64 mesh
    // The mesh's own copies of a sorted slot order (slot -> atom, -1
    // padding), two halves of nslots: a step's spread and gather read half
    // order_sel while the pair kernels' fold copies the list's order into the
    // other half (pme_next_order), which the next step reads.  The mesh may
    // then run while a list build rewrites nl.sorted_atom and still visit
    // the atoms in a spatial order.  Both halves hold orders of the list
    // layout of generation order_gen (any such order gives the same sums).
    DevBuf<int> order;
    int order_gen = -1;
    int order_sel = 0;
};

// Grid dimensions must each be >= 5 (pme.rs:222-225) and at most 1024.
// Reads sys[s].grid and sys[s].alpha; writes the grid offsets, radices,
// root/moduli offsets and FFT line counts into sys[s].
void pme_setup(PmeDev& pm, std::vector<SysInfo>& sys, int natoms, const int* atom_sys, bool mixed, hipStream_t s);

// FFT workgroup grouping of precision prec (0: FP32, 1: FP64): at most cap
// lines per workgroup (cap <= 0: the default rule); writes sys[s].fft_lines,
// which the caller uploads.  Returns the cap.  Results do not depend on it.
int pme_set_fft_lines(PmeDev& pm, std::vector<SysInfo>& sys, int prec, int cap, hipStream_t s);
// FFT decomposition of precision prec: use = 1 the plane passes where
// available, 0 the five line passes, < 0 the default (planes where
// available).  Writes sys[s].plane_lines, which the caller uploads; returns
// whether the plane passes are in use.  Results do not depend on it.
bool pme_set_fft_plane(PmeDev& pm, std::vector<SysInfo>& sys, int prec, int use);
// Spread tile edge (6 or 8; <= 0: the default rule), with tiles and bins to
// match; writes sys[s].ptile/tiles/tile_off.  Results do not depend on it.
int pme_set_spread_edge(PmeDev& pm, std::vector<SysInfo>& sys, int edge, hipStream_t s);

// Doubles the per-tile bin capacity after ST_BIN_OVERFLOW.
void pme_grow_bins(PmeDev& pm);
// Reallocates the bins to `cap` atoms per spread tile.
void pm_set_bin_cap(PmeDev& pm, int cap);
// Copies the list's sorted order into both halves of the mesh's own order
// (after a list build outside the step loop; allocates).
void pme_take_order(PmeDev& pm, const NlistDev& nl, hipStream_t s);
// The same copy into pm's and, when given, other's orders by a kernel
// (graph-capturable: allocates nothing; skips an order not allocated for
// this layout's size).  s must follow both the list build and every mesh
// phase that reads the orders.
void pme_follow_order(PmeDev& pm, PmeDev* other, const NlistDev& nl, hipStream_t s);
// Whether pm holds orders of the current list layout.
bool pme_order_ready(const PmeDev& pm, const NlistDev& nl);
// After a step's mesh phases have been enqueued: the half of pm's order they
// do not read (where the step's list order is to be copied, on a stream that
// follows the list build), which the next step's mesh then reads.
int* pme_next_order(PmeDev& pm, const NlistDev& nl);

// Spread, FFT, convolution (+energy into slot E_RECIP of each system), inverse
// FFT, gather (forces added).  Atoms with zero charge are skipped as in the
// reference.
enum : int { PME_SPREAD = 1, PME_FFT = 2, PME_GATHER = 4, PME_ALL = 7 };
// Arithmetic of one PME evaluation.
//   MESH_F32      : FP32 splines/FFT/gather, fixed-point grid and forces (mixed
//                   form, force-only MD steps; no mesh energy is summed);
//   MESH_F64_FIXED: FP64 splines/FFT/gather on the same fixed-point grid with
//                   fixed-point forces (mixed form when the energy is wanted:
//                   the mesh energy is ~1e4-1e5 kJ/mol and must carry ~1e-9
//                   relative error);
//   MESH_F64      : FP64 throughout with FP64 atomics (plain form).
enum MeshMode : int { MESH_F32 = 0, MESH_F64_FIXED = 1, MESH_F64 = 2 };
// The atoms are visited in the list's sorted slot order (nl.sorted_atom),
// which only sets locality: any permutation gives the same fixed-point sums.
// atom_order: the spread and the gather do not read nl.sorted_atom but the
// mesh's own order (the sorted order of an earlier build of the same layout,
// copied by pme_take_order or pme_follow_order), so the mesh may run while a
// list build rewrites the sorted order.  Without a valid own order the spread
// takes index order and the gather nl.sorted_atom (it must then follow the
// build).
void pme_run(PmeDev& pm, int mode, const NlistDev& nl, const double* pos, const double* charge,
             const SysInfo* sys_dev, long long* force_fx, double* force_d, long long* energy_fx, double* energy_d,
             unsigned* status, hipStream_t s, int phases = PME_ALL, bool atom_order = false);

// out[k] = the FP64 mesh potential at the stencil of atom atoms[k] of
// system 0 (dE/dq of that atom, pme.rs:117-137), after an FP64 mesh run
// (MESH_F64_FIXED or MESH_F64) of the spread and FFT phases.
void pme_potential_at(const PmeDev& pm, const SysInfo* sys_dev, const double* pos, const int* atoms, int count,
                      double* out, hipStream_t s);

// B-spline moduli |sum_i M5(i+1) e^{2 pi i k i/n}|^2 with near-zero values
// replaced by the mean of their neighbours (pme.rs:239-259).
std::vector<double> bspline_moduli(int n);

}  // namespace ommhip
