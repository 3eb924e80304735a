// Device-resident MD step for the flexible-water workloads: NonbondedForce
// (PME) + HarmonicBondForce + HarmonicAngleForce + LangevinMiddle, captured
// into a HIP graph (Context::step, context.rs:795-928).
#include <cfloat>
#include <climits>
#include <cmath>
#include <functional>
#include <random>
#include <map>
#include <numeric>
#include <set>
#include <type_traits>

#include "md.h"
#include "terms.h"

namespace ommhip {

namespace {

// BOLTZMANN in kJ/mol/K (langevin.rs:12).
const double BOLTZMANN = 1.380649e-23 * 6.02214076e23 / 1000.0;

// Bonded terms in one launch.  Each small molecule (at most four atoms
// joined by bonds, angles or exceptions, e.g. a water) takes GROUP_LANES
// consecutive lanes: lane k evaluates the molecule's terms k, k + GROUP_LANES,
// ... (its exceptions, then bonds, then angles) into per-atom sums in
// registers, the lanes' sums are combined by a fixed butterfly of shuffles,
// and the molecule's atoms receive one fixed-point add per component.  The
// threads after the molecules' lanes take the terms outside such molecules
// one each.  Indices are global over the batch; energies go to the term's
// system.
constexpr int GROUP_LANES = 8;
__device__ __forceinline__ int group_slot(const int at[4], int i) {
    return i == at[0] ? 0 : (i == at[1] ? 1 : (i == at[2] ? 2 : 3));
}

// f[slot] += v, with the slot chosen by selects (a register array; indexing
// it with a run-time slot would put it in scratch memory).
__device__ __forceinline__ void add_slot(double f[4][3], int slot, const double v[3]) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) f[k][c] += slot == k ? v[c] : 0.0;
}

template <class A>
__global__ __launch_bounds__(256) void k_bonded(
    int ngroups, const int* __restrict__ grp_atoms, const int* __restrict__ grp_e, const int* __restrict__ grp_b,
    const int* __restrict__ grp_a, int ne, int ge, const int* __restrict__ e1, const int* __restrict__ e2,
    const double* __restrict__ eqq, const double* __restrict__ esig, const double* __restrict__ eeps, int nb, int gb,
    const int* __restrict__ b1, const int* __restrict__ b2, const double* __restrict__ blen,
    const double* __restrict__ bk, int na, int ga, const int* __restrict__ a1, const int* __restrict__ a2,
    const int* __restrict__ a3, const double* __restrict__ ath, const double* __restrict__ ak,
    const double* __restrict__ charge, int use_pbc, int has_pme, double alpha, const double* __restrict__ ec6,
    double lj_alpha, const double* __restrict__ pos,
    const int* __restrict__ atom_sys, const SysInfo* __restrict__ sys, A* force, A* energy, unsigned* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double ee = 0.0, eb = 0.0, ea = 0.0;
    int se = -1, sb = -1, sa = -1;
    const int grouped = ngroups * GROUP_LANES;  // a multiple of GROUP_LANES: a molecule's lanes share a branch
    if (t < grouped) {
        const int g = t / GROUP_LANES, lane = t % GROUP_LANES;
        int at[4];
        for (int k = 0; k < 4; ++k) at[k] = grp_atoms[4 * g + k];
        const int si = atom_sys[at[0]];
        const Box& box = sys[si].box;
        const int e0 = grp_e[g], ne_g = grp_e[g + 1] - e0;
        const int b0 = grp_b[g], nb_g = grp_b[g + 1] - b0;
        const int a0 = grp_a[g], na_g = grp_a[g + 1] - a0;
        double f[4][3] = {};
        for (int k = lane; k < ne_g + nb_g + na_g; k += GROUP_LANES) {
            if (k < ne_g) {
                const int e = e0 + k, i = e1[e], j = e2[e];
                double fi[3];
                ee += exception_pair(i, j, eqq[e], esig[e], eeps[e], charge[i], charge[j], pos, box, use_pbc,
                                     has_pme, alpha, ec6 ? ec6[e] : 0.0, lj_alpha, fi);
                const double mi[3] = {-fi[0], -fi[1], -fi[2]};
                add_slot(f, group_slot(at, i), fi);
                add_slot(f, group_slot(at, j), mi);
            } else if (k < ne_g + nb_g) {
                const int b = b0 + k - ne_g, i = b1[b], j = b2[b];
                double fi[3];
                eb += bond_pair(i, j, blen[b], bk[b], pos, fi);
                const double mi[3] = {-fi[0], -fi[1], -fi[2]};
                add_slot(f, group_slot(at, i), fi);
                add_slot(f, group_slot(at, j), mi);
            } else {
                const int a = a0 + k - ne_g - nb_g, i = a1[a], j = a2[a], l = a3[a];
                double fi[3], fk[3];
                ea += angle_triple(i, j, l, ath[a], ak[a], pos, fi, fk);
                const double fj[3] = {-(fi[0] + fk[0]), -(fi[1] + fk[1]), -(fi[2] + fk[2])};
                add_slot(f, group_slot(at, i), fi);
                add_slot(f, group_slot(at, l), fk);
                add_slot(f, group_slot(at, j), fj);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int o = 1; o < GROUP_LANES; o <<= 1) f[k][c] += __shfl_xor(f[k][c], o, 64);
        // Lane k adds the components k and k + GROUP_LANES of the 4 x 3.
#pragma unroll
        for (int q = 0; q < 12; q += GROUP_LANES) {
            const int kc = q + lane;
            if (kc < 12 && at[kc / 3] >= 0) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int c = 0; c < 3; ++c) v = kc == 3 * k + c ? f[k][c] : v;
                acc_add(force + 3 * at[kc / 3] + kc % 3, v, FORCE_SCALE, status);
            }
        }
        se = ne_g > 0 ? si : -1;
        sb = nb_g > 0 ? si : -1;
        sa = na_g > 0 ? si : -1;
    } else {
        int u = t - grouped;
        const int le = ne - ge, lb = nb - gb, la = na - ga;
        if (u < le) {
            const int e = ge + u, i = e1[e], j = e2[e];
            se = atom_sys[i];
            ee = exception_term(i, j, eqq[e], esig[e], eeps[e], charge[i], charge[j], pos, sys[se].box, use_pbc,
                                has_pme, alpha, ec6 ? ec6[e] : 0.0, lj_alpha, force, status);
        } else if ((u -= le) < lb) {
            const int b = gb + u;
            sb = atom_sys[b1[b]];
            eb = bond_term(b1[b], b2[b], blen[b], bk[b], pos, force, status);
        } else if ((u -= lb) < la) {
            const int a = ga + u;
            sa = atom_sys[a1[a]];
            ea = angle_term(a1[a], a2[a], a3[a], ath[a], ak[a], pos, force, status);
        }
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

// Last kernel of a run of `count` steps (one eager step, or a captured
// graph): advance the step counter, clear the energies.  Inside the run, step
// k's integrator draws at counter step[0] + k + 1, so a graph needs one such
// node rather than one per step.
template <class A>
__global__ void k_step_advance(unsigned long long* step, int count, A* energy, int nslots) {
    if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += (unsigned long long)count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += gridDim.x * blockDim.x) energy[i] = (A)0;
}

// Snapshot of the state a run of steps starts from: positions, velocities,
// the positions the neighbour list was built from (nullable when there is no
// list) and the step counter.
__global__ void k_snapshot(const double* __restrict__ pos, double* __restrict__ pos_out,
                           const double* __restrict__ vel, double* __restrict__ vel_out,
                           const double* __restrict__ built, double* __restrict__ built_out, size_t n,
                           const unsigned long long* __restrict__ step, unsigned long long* step_out) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) {
        pos_out[i] = pos[i];
        vel_out[i] = vel[i];
        if (built) built_out[i] = built[i];
    }
    if (i == 0) step_out[0] = step[0];
}

inline __device__ double force_value(long long f) { return (double)f * (1.0 / FORCE_SCALE); }
inline __device__ double force_value(double f) { return f; }

// The three standard normals of atom i (global index) at the current step:
// Philox keyed by the system's seed, counter (index within the system,
// step), then Box-Muller in FP32 with the hardware log2 and sin/cos (whose
// argument is in turns): the normals only scale the noise term, and ~1e-7
// relative error in them is statistically invisible.
__device__ __forceinline__ void step_normals(int i, const SysInfo& S, unsigned long long step, double xi[3]) {
    const uint64_t seed = S.seed;
    uint32_t c[4] = {(uint32_t)(i - S.atom_off), (uint32_t)step, (uint32_t)(step >> 32), 0x4c464d31u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u1 = ((float)(c[0] >> 8) + 1.0f) * (1.0f / 16777216.0f);  // (0, 1]
    const float u2 = (float)(c[1] >> 8) * (1.0f / 16777216.0f);           // [0, 1)
    const float u3 = ((float)(c[2] >> 8) + 1.0f) * (1.0f / 16777216.0f);
    const float u4 = (float)(c[3] >> 8) * (1.0f / 16777216.0f);
    const float ln2 = 0.6931471805599453f;
    const float r1 = __builtin_sqrtf(-2.0f * ln2 * __builtin_amdgcn_logf(u1));
    const float r2 = __builtin_sqrtf(-2.0f * ln2 * __builtin_amdgcn_logf(u3));
    xi[0] = (double)(r1 * __builtin_amdgcn_cosf(u2));
    xi[1] = (double)(r1 * __builtin_amdgcn_sinf(u2));
    xi[2] = (double)(r2 * __builtin_amdgcn_cosf(u4));
}

// LangevinMiddle update (langevin.rs:103-142) with counter-based normals:
// Philox keyed by the system's seed, counter (index within the system,
// step), so a replica draws the same numbers alone or in a batch.  The force
// accumulator is read and cleared for the next step.  check != 0: raise the
// neighbour list's rebuild flags (any, and the atom's system) when an atom
// moved more than skin/2 from its position at the last build, so the next
// step's build kernels rebuild that system.
// A non-periodic system's atom outside the region its virtual box serves.
__device__ __forceinline__ bool left_region(const SysInfo& S, const double* x) {
    if (!S.vcheck) return false;
    bool out = false;
    for (int a = 0; a < 3; ++a) out = out || !(x[a] >= S.vlo[a] && x[a] <= S.vhi[a]);
    return out;
}

template <class A>
__global__ void k_lfmiddle(int n, double* __restrict__ pos, double* __restrict__ vel, A* __restrict__ force,
                           const double* __restrict__ inv_mass, const int* __restrict__ atom_sys,
                           const SysInfo* __restrict__ sys, double dt, double decay, double variance,
                           const unsigned long long* __restrict__ step, int step_offset,
                           const double* __restrict__ pos_build, double lim2, int check, int* flag,
                           const unsigned char* __restrict__ constrained, unsigned* status) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false, outside = false;
    int si = 0;
    if (i < n && !(constrained && constrained[i])) {  // constrained atoms: k_constrained_step
        double im = inv_mass[i];
        double f[3];
        for (int a = 0; a < 3; ++a) {
            f[a] = force_value(force[3 * i + a]);
            force[3 * i + a] = (A)0;
        }
        if (im != 0.0) {  // massless particles do not move
            si = atom_sys[i];
            double xi[3];
            step_normals(i, sys[si], step[0] + (unsigned long long)step_offset, xi);
            double noise = sqrt(variance * im);
            double half = 0.5 * dt;
            double d[3];
            for (int a = 0; a < 3; ++a) {
                int k = 3 * i + a;
                double v = vel[k] + f[a] * im * dt;
                double mid = pos[k] + half * v;
                v = decay * v + noise * xi[a];
                vel[k] = v;
                double x = mid + half * v;
                pos[k] = x;
                d[a] = x - pos_build[k];
            }
            moved = check && !(skin_d2(d[0], d[1], d[2]) <= lim2);
            outside = left_region(sys[si], pos + 3 * i);
        }
    }
    wave_raise_flags(flag, moved, si);
    if (__ballot(outside) && (threadIdx.x & 63) == 0) atomicOr(status, ST_VBOX);
}

// One SHAKE (velocities = false) or RATTLE (velocities = true) projection of
// a component's constraints (constraints.rs:80-149): Gauss-Seidel sweeps in
// the reference's order, directions from the reference positions `ref`,
// until every constraint is within tolerance (at most 1000 sweeps).  Returns
// false on the reference's errors (singular or nonfinite projection, no
// convergence).
__device__ bool project_component(int k0, int k1, const int* __restrict__ c1, const int* __restrict__ c2,
                                  const double* __restrict__ d2, const double* __restrict__ w1,
                                  const double* __restrict__ w2, const double* __restrict__ ref, double* values,
                                  bool velocities, double tolerance) {
    for (int sweep = 0; sweep < 1000; ++sweep) {
        bool converged = true;
        for (int k = k0; k < k1; ++k) {
            const int i = c1[k], j = c2[k];
            const double wi = w1[k], wj = w2[k];
            double dir[3], delta[3], dot = 0.0;
            for (int a = 0; a < 3; ++a) {
                dir[a] = ref[3 * i + a] - ref[3 * j + a];
                delta[a] = values[3 * i + a] - values[3 * j + a];
                dot += delta[a] * dir[a];
            }
            double mult;
            if (velocities) {
                const double norm = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
                if (norm == 0.0 || !isfinite(norm)) return false;
                const double correction = -dot / ((wi + wj) * norm);
                if (fabs(correction) <= tolerance) continue;
                mult = correction;
            } else {
                const double norm = delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2];
                const double ratio = sqrt(norm / d2[k]);
                if (isfinite(ratio) && fabs(ratio - 1.0) <= tolerance) continue;
                const double denominator = 2.0 * (wi + wj) * dot;
                if (fabs(denominator) < DBL_MIN || !isfinite(denominator)) return false;
                mult = (d2[k] - norm) / denominator;
            }
            if (!isfinite(mult)) return false;
            converged = false;
            for (int a = 0; a < 3; ++a) {
                const double correction = mult * dir[a];
                values[3 * i + a] += wi * correction;
                values[3 * j + a] -= wj * correction;
            }
        }
        if (converged) return true;
    }
    return false;
}

// Constrained LangevinMiddle (langevin.rs:215-277), one thread per
// constraint component: kick, RATTLE the velocities against the current
// positions, drift / friction + noise (the same normals as k_lfmiddle) /
// drift, SHAKE the positions against the previous ones, and add the position
// correction / dt to the velocities.  mode 1 projects only the positions
// (against themselves), mode 2 only the velocities (against the positions):
// Context::apply_constraints / apply_velocity_constraints.
template <class A>
__global__ void k_constrained_step(int ncomp, const int* __restrict__ atom_off, const int* __restrict__ atoms,
                                   const int* __restrict__ cons_off, const int* __restrict__ c1,
                                   const int* __restrict__ c2, const double* __restrict__ d2,
                                   const double* __restrict__ w1, const double* __restrict__ w2, double tolerance,
                                   int mode, double* pos, double* vel, A* __restrict__ force,
                                   const double* __restrict__ inv_mass, const int* __restrict__ atom_sys,
                                   const SysInfo* __restrict__ sys, double dt, double decay, double variance,
                                   const unsigned long long* __restrict__ step, int step_offset,
                                   const double* __restrict__ pos_build, double lim2, int check, int* flag,
                                   double* prev, double* proposed, unsigned* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    int si = 0;
    if (t < ncomp) {
        const int a0 = atom_off[t], a1 = atom_off[t + 1], k0 = cons_off[t], k1 = cons_off[t + 1];
        si = atom_sys[atoms[a0]];
        bool ok = true;
        if (mode == 2) {
            ok = project_component(k0, k1, c1, c2, d2, w1, w2, pos, vel, true, tolerance);
        } else if (mode == 1) {
            for (int m = a0; m < a1; ++m)
                for (int a = 0; a < 3; ++a) prev[3 * atoms[m] + a] = pos[3 * atoms[m] + a];
            ok = project_component(k0, k1, c1, c2, d2, w1, w2, prev, pos, false, tolerance);
        } else {
            for (int m = a0; m < a1; ++m) {
                const int i = atoms[m];
                const double im = inv_mass[i];
                for (int a = 0; a < 3; ++a) {
                    prev[3 * i + a] = pos[3 * i + a];
                    const double f = force_value(force[3 * i + a]);
                    force[3 * i + a] = (A)0;
                    vel[3 * i + a] += f * im * dt;
                }
            }
            ok = project_component(k0, k1, c1, c2, d2, w1, w2, prev, vel, true, tolerance);
            const double half = 0.5 * dt;
            for (int m = a0; m < a1; ++m) {
                const int i = atoms[m];
                double xi[3];
                step_normals(i, sys[si], step[0] + (unsigned long long)step_offset, xi);
                const double noise = sqrt(variance * inv_mass[i]);
                for (int a = 0; a < 3; ++a) {
                    const int k = 3 * i + a;
                    double v = vel[k];
                    const double mid = pos[k] + half * v;
                    v = decay * v + noise * xi[a];
                    vel[k] = v;
                    pos[k] = mid + half * v;
                    proposed[k] = pos[k];
                }
            }
            ok = ok && project_component(k0, k1, c1, c2, d2, w1, w2, prev, pos, false, tolerance);
            bool outside = false;
            for (int m = a0; m < a1; ++m) {
                const int i = atoms[m];
                double d[3];
                for (int a = 0; a < 3; ++a) {
                    const int k = 3 * i + a;
                    vel[k] += (pos[k] - proposed[k]) / dt;
                    d[a] = pos[k] - pos_build[k];
                }
                moved = moved || (check && !(skin_d2(d[0], d[1], d[2]) <= lim2));
                outside = outside || left_region(sys[si], pos + 3 * i);
            }
            if (outside) atomicOr(status, ST_VBOX);
        }
        if (!ok) atomicOr(status, ST_CONSTRAINTS);
    }
    if (mode == 0) wave_raise_flags(flag, moved, si);
}

// Monte Carlo barostat (barostat.rs:129-195, context/program.rs:5-45,
// context.rs:976-1029), one thread per system or molecule.
//
// Proposal: V = a_x b_y c_z, dV = volume_scale * 2 (u - 1/2), V' = V + dV; an
// impossible V' counts as a rejected attempt without a move.  The box is
// scaled by s = (V'/V)^(1/3) and every molecule is translated by (s - 1) times
// its unweighted centroid.  The two uniforms come from Philox keyed by the
// barostat seed + replica index, counter = attempt number, so a replica draws
// the same numbers alone or in a batch.
//
// The anisotropic form (anisotropic_barostat.rs:122-211) draws one of the
// enabled axes (a third uniform, uniform over them), scales that axis by
// V'/V, and keeps the step size and counters per axis.
__device__ __forceinline__ void adapt_step(double& volume_scale, long long& attempted, long long& accepted_count,
                                           bool accepted, double previous_volume) {
    // barostat.rs:180-195, anisotropic_barostat.rs:190-211
    attempted += 1;
    accepted_count += accepted ? 1 : 0;
    if (attempted >= 10) {
        const double fraction = (double)accepted_count / (double)attempted;
        if (fraction < 0.25) {
            volume_scale /= 1.1;
        } else if (fraction > 0.75) {
            volume_scale = fmin(volume_scale * 1.1, previous_volume * 0.3);
        } else {
            return;
        }
        attempted = 0;
        accepted_count = 0;
    }
}

__device__ __forceinline__ void baro_record(BaroState& B, int kind, bool accepted, double previous_volume) {
    if (kind == 3) {
        // flexible_barostat.rs:200-215: a length step, capped at 0.3 V^(1/3).
        B.attempted += 1;
        B.accepted += accepted ? 1 : 0;
        if (B.attempted >= 10) {
            const double fraction = (double)B.accepted / (double)B.attempted;
            if (fraction < 0.25) {
                B.volume_scale /= 1.1;
            } else if (fraction > 0.75) {
                B.volume_scale = fmin(B.volume_scale * 1.1, cbrt(previous_volume) * 0.3);
            } else {
                return;
            }
            B.attempted = 0;
            B.accepted = 0;
        }
    } else if (kind != 0) {
        const int a = B.axis;
        adapt_step(B.axis_volume_scale[a], B.axis_attempted[a], B.axis_accepted[a], accepted, previous_volume);
        B.volume_scale = B.axis_volume_scale[a];
        B.attempted = B.axis_attempted[a];
        B.accepted = B.axis_accepted[a];
    } else {
        adapt_step(B.volume_scale, B.attempted, B.accepted, accepted, previous_volume);
    }
}

// flexible_barostat.rs:15-41: the box vectors t (row-major) put in reduced
// form in place: each diagonal element made positive by flipping its vector,
// then c -= round(c_y / b_y) b, c -= round(c_x / a_x) a, b -= round(b_x / a_x) a.
// False when an element is not finite or a diagonal element is zero.
__host__ __device__ inline bool reduce_box(double t[9]) {
#pragma clang fp contract(off)
    bool valid = true;
    for (int k = 0; k < 9; ++k) valid = valid && __builtin_isfinite(t[k]);
    for (int row = 0; row < 3 && valid; ++row) {
        if (t[3 * row + row] == 0.0) valid = false;
        else if (t[3 * row + row] < 0.0)
            for (int col = 0; col < 3; ++col) t[3 * row + col] = -t[3 * row + col];
    }
    const int red[3][3] = {{2, 1, 1}, {2, 0, 0}, {1, 0, 0}};  // (target, source, component)
    for (int k = 0; k < 3 && valid; ++k) {
        const int tg = red[k][0], sr = red[k][1], cp = red[k][2];
        const double shift = round(t[3 * tg + cp] / t[3 * sr + cp]);  // half away from zero, as f64::round
        for (int col = 0; col < 3; ++col) t[3 * tg + col] -= shift * t[3 * sr + col];
    }
    for (int k = 0; k < 9; ++k) valid = valid && __builtin_isfinite(t[k]);
    return valid;
}

__global__ void k_baro_propose(int nsys, SysInfo* sys, BaroState* st, Box* saved_box, double* scale, uint64_t seed,
                               double cutoff, int kind, int scale_axes, int xy_mode, int z_mode, unsigned* status) {
    const bool anisotropic = kind != 0;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsys) return;
    BaroState& B = st[r];
    Box& box = sys[r].box;
    saved_box[r] = box;
    const uint64_t key = seed + (uint64_t)r;
    uint32_t c[4] = {(uint32_t)B.total_attempted, (uint32_t)((uint64_t)B.total_attempted >> 32), 0u, 0x42415230u};
    philox4x32_10(c, (uint32_t)key, (uint32_t)(key >> 32));
    B.u_volume = (double)c[0] * (1.0 / 4294967296.0);
    B.u_accept = (double)c[1] * (1.0 / 4294967296.0);
    B.total_attempted += 1;
    double step = B.volume_scale;
    const double u_axis = (double)c[2] * (1.0 / 4294967296.0);
    if (kind == 2) {
        // membrane_barostat.rs:178-190: x or y (one axis when coupled), or z
        // when it is free, each draw of the three uniform.
        int axis;
        if (z_mode == 0) {
            const int cand = min((int)(u_axis * 3.0), 2);
            axis = cand < 2 ? (xy_mode == 0 ? 0 : cand) : 2;
        } else {
            const int cand = min((int)(u_axis * 2.0), 1);
            axis = xy_mode == 0 ? 0 : cand;
        }
        B.axis = axis;
        step = B.axis_volume_scale[axis];
    } else if (anisotropic) {
        // The k-th enabled axis, k uniform (anisotropic_barostat.rs:133-139).
        const int enabled = __popc(scale_axes & 7);
        int k = (int)(u_axis * enabled);
        if (k >= enabled) k = enabled - 1;
        int axis = 0;
        for (int a = 0; a < 3; ++a)
            if ((scale_axes >> a) & 1) {
                if (k == 0) {
                    axis = a;
                    break;
                }
                --k;
            }
        B.axis = axis;
        step = B.axis_volume_scale[axis];
    }
    const double v_old = box.v[0] * box.v[4] * box.v[8];
    double v_new = v_old + step * 2.0 * (B.u_volume - 0.5);
    B.v_old = v_old;
    B.v_new = v_new;
    B.accepted_last = 0;
    B.area_change = 0.0;
    B.log_shape = 0.0;
    for (int a = 0; a < 3; ++a) B.scale[a] = scale[3 * r + a] = 1.0;
    if (kind == 3) {
        // flexible_barostat.rs:139-168: every lower-triangle element moves by
        // up to the length step; the trial box is put back in reduced form.
        uint32_t c2[4] = {(uint32_t)B.total_attempted, (uint32_t)((uint64_t)B.total_attempted >> 32), 1u,
                          0x42415230u};
        philox4x32_10(c2, (uint32_t)key, (uint32_t)(key >> 32));
        const double u[6] = {B.u_volume, u_axis, (double)c[3] * (1.0 / 4294967296.0),
                             (double)c2[0] * (1.0 / 4294967296.0), (double)c2[1] * (1.0 / 4294967296.0),
                             (double)c2[2] * (1.0 / 4294967296.0)};
        double t[9];
        for (int k = 0; k < 9; ++k) t[k] = box.v[k];
        int q = 0;
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col <= row; ++col) t[3 * row + col] += B.volume_scale * 2.0 * (u[q++] - 0.5);
        const bool valid = reduce_box(t);
        v_new = valid ? t[0] * t[4] * t[8] : NAN;
        B.v_new = v_new;
        for (int k = 0; k < 9; ++k) B.box_new[k] = valid ? t[k] : box.v[k];
        if (!(isfinite(v_new) && v_new > 0.0)) {
            B.skipped = 1;
            baro_record(B, kind, false, v_old);
            return;
        }
        B.skipped = 0;
        double sc[3];
        for (int a = 0; a < 3; ++a) sc[a] = t[4 * a] / box.v[4 * a];
        B.log_shape = 2.0 * log(sc[0]) + log(sc[1]);
        if (!(2.0 * cutoff <= t[0] && 2.0 * cutoff <= t[4] && 2.0 * cutoff <= t[8])) {
            atomicOr(status, ST_CUTOFF_BOX);
            B.skipped = 1;
            return;
        }
        for (int a = 0; a < 3; ++a) B.scale[a] = scale[3 * r + a] = sc[a];
        for (int k = 0; k < 9; ++k) box.v[k] = t[k];
        box_refresh(box);
        return;
    }
    if (!(isfinite(v_new) && v_new > 0.0)) {
        B.skipped = 1;
        baro_record(B, kind, false, v_old);
        return;
    }
    B.skipped = 0;
    double sc[3];
    if (kind == 2) {
        // membrane_barostat.rs:198-209
        sc[0] = sc[1] = sc[2] = 1.0;
        if (B.axis < 2 && xy_mode == 0) {
            sc[0] = sc[1] = sqrt(v_new / v_old);
        } else {
            sc[B.axis] = v_new / v_old;
        }
        if (z_mode == 2) {
            sc[2] = 1.0 / (sc[0] * sc[1]);
            v_new = v_old;
            B.v_new = v_new;
        }
        B.area_change = box.v[0] * box.v[4] * (sc[0] * sc[1] - 1.0);
    } else if (anisotropic) {
        for (int a = 0; a < 3; ++a) sc[a] = a == B.axis ? v_new / v_old : 1.0;
    } else {
        sc[0] = sc[1] = sc[2] = cbrt(v_new / v_old);
    }
    // The reference asserts 2 rc <= box edge on every evaluation
    // (nonbonded.rs:806-812); a move that would break it is an error.
    if (!(2.0 * cutoff <= box.v[0] * sc[0] && 2.0 * cutoff <= box.v[4] * sc[1] && 2.0 * cutoff <= box.v[8] * sc[2])) {
        atomicOr(status, ST_CUTOFF_BOX);
        B.skipped = 1;
        return;
    }
    for (int a = 0; a < 3; ++a) B.scale[a] = scale[3 * r + a] = sc[a];
    // barostat.rs:28-40: each box vector's component a scaled by sc[a].
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) box.v[3 * k + a] *= sc[a];
    box_refresh(box);
}

__global__ void k_baro_scale(int nmol, const int* __restrict__ mol_start, const int* __restrict__ mol_atoms,
                             const int* __restrict__ mol_sys, const double* __restrict__ scale, double* pos,
                             double* pos_saved) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int lo = mol_start[m], hi = mol_start[m + 1];
    double center[3] = {0.0, 0.0, 0.0};
    for (int k = lo; k < hi; ++k) {
        const int a = mol_atoms[k];
        for (int d = 0; d < 3; ++d) {
            pos_saved[3 * a + d] = pos[3 * a + d];
            center[d] += pos[3 * a + d];
        }
    }
    const double* sc = scale + 3 * mol_sys[m];
    if (sc[0] == 1.0 && sc[1] == 1.0 && sc[2] == 1.0) return;
    const double inv = 1.0 / (double)(hi - lo);
    double offset[3];
    for (int d = 0; d < 3; ++d) offset[d] = (sc[d] - 1.0) * (center[d] * inv);
    for (int k = lo; k < hi; ++k) {
        const int a = mol_atoms[k];
        for (int d = 0; d < 3; ++d) pos[3 * a + d] += offset[d];
    }
}

inline __device__ double energy_value(long long e) { return (double)e * (1.0 / ENERGY_SCALE); }
inline __device__ double energy_value(double e) { return e; }

// Potential energy of one system from its energy slots plus the volume terms.
template <class A>
__device__ double system_energy(const A* e, const SysInfo& S, double volume) {
    return energy_value(e[E_DIRECT]) + energy_value(e[E_RECIP]) + S.self_const +
           (S.neutral_coeff + S.disp_coeff) / volume + energy_value(e[E_EXCEPT]) + energy_value(e[E_BOND]) +
           energy_value(e[E_ANGLE]);
}

template <class A>
__global__ void k_baro_accept(int nsys, SysInfo* sys, BaroState* st, const Box* __restrict__ saved_box, int* rejected,
                              const A* __restrict__ e_old, const A* __restrict__ e_new,
                              const int* __restrict__ nmol_sys, double pressure, int kind, double px, double py,
                              double pz, double surface_tension, double kt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsys) return;
    BaroState& B = st[r];
    if (B.skipped) {
        rejected[r] = 0;  // nothing moved
        return;
    }
    SysInfo& S = sys[r];
    const double eo = system_energy(e_old + (size_t)r * E_SLOTS, S, saved_box[r].volume);
    const double en = system_energy(e_new + (size_t)r * E_SLOTS, S, S.box.volume);
    // barostat.rs:160-178, anisotropic_barostat.rs:162-188 (BAR_NM3_TO_KJ_PER_MOL = 0.0602214076)
    // membrane_barostat.rs:229-233: - surface tension x area change;
    // flexible_barostat.rs:177-183: (N - 2) kT ln(V'/V) and - kT x shape.
    if (kind == 1) pressure = B.axis == 0 ? px : (B.axis == 1 ? py : pz);
    const double nmol = (double)nmol_sys[r] - (kind == 3 ? 2.0 : 0.0);
    const double work = (en - eo) + pressure * 0.0602214076 * (B.v_new - B.v_old) -
                        (kind == 2 ? surface_tension * 0.0602214076 * B.area_change : 0.0) -
                        nmol * kt * log(B.v_new / B.v_old) - (kind == 3 ? kt * B.log_shape : 0.0);
    bool accepted;
    if (isnan(work)) accepted = false;
    else if (work <= 0.0) accepted = true;
    else accepted = kt > 0.0 && B.u_accept <= exp(-work / kt);
    B.e_old = eo;
    B.e_new = en;
    B.work = work;
    B.accepted_last = accepted ? 1 : 0;
    B.total_accepted += accepted ? 1 : 0;
    baro_record(B, kind, accepted, kind == 3 ? B.v_new : B.v_old);
    if (!accepted) S.box = saved_box[r];
    rejected[r] = accepted ? 0 : 1;
}

// Rigid-molecule move by a per-system 3x3 transform T (row-major): every atom
// of a molecule is shifted by (T - I) c, c the molecule's geometric centre
// (context.rs:976-1029 with rigid molecules).  A diagonal T multiplies the
// diagonal only; a general T sums each row in order, without contraction.
__global__ void k_transform_molecules(int nmol, const int* __restrict__ mol_start, const int* __restrict__ mol_atoms,
                                      const int* __restrict__ mol_sys, const double* __restrict__ transform,
                                      double* pos) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int lo = mol_start[m], hi = mol_start[m + 1];
    double c[3] = {0.0, 0.0, 0.0};
    for (int k = lo; k < hi; ++k)
        for (int d = 0; d < 3; ++d) c[d] += pos[3 * mol_atoms[k] + d];
    const double inv = 1.0 / (double)(hi - lo);
    for (int d = 0; d < 3; ++d) c[d] *= inv;
    double t[9];
    for (int k = 0; k < 9; ++k) t[k] = transform[9 * (size_t)mol_sys[m] + k];
    for (int d = 0; d < 3; ++d) t[4 * d] -= 1.0;
    const bool diagonal = t[1] == 0.0 && t[2] == 0.0 && t[3] == 0.0 && t[5] == 0.0 && t[6] == 0.0 && t[7] == 0.0;
    double off[3];
    for (int d = 0; d < 3; ++d)
        off[d] = diagonal ? t[4 * d] * c[d] : (t[3 * d] * c[0] + t[3 * d + 1] * c[1]) + t[3 * d + 2] * c[2];
    for (int k = lo; k < hi; ++k)
        for (int d = 0; d < 3; ++d) pos[3 * mol_atoms[k] + d] += off[d];
}

// Trial boxes of the pressure queries: with `boxes`, each system's box is
// saved and replaced by boxes[9 r ..]; without, the saved box is put back.
__global__ void k_swap_boxes(int nsys, SysInfo* sys, Box* saved, const double* __restrict__ boxes) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsys) return;
    if (boxes) {
        saved[r] = sys[r].box;
        Box b = sys[r].box;
        for (int k = 0; k < 9; ++k) b.v[k] = boxes[9 * (size_t)r + k];
        box_refresh(b);
        sys[r].box = b;
    } else {
        sys[r].box = saved[r];
    }
}

__global__ void k_baro_restore(int n, const int* __restrict__ atom_sys, const int* __restrict__ rejected, double* pos,
                               const double* __restrict__ pos_saved) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !rejected[atom_sys[i]]) return;
    for (int d = 0; d < 3; ++d) pos[3 * i + d] = pos_saved[3 * i + d];
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
    if (nb_->is_virtual()) {
        // NoCutoff / CutoffNonPeriodic: the periodic machinery on a virtual
        // box, set up here for all atoms at the origin and chosen again from
        // the positions when they are set (prepare).  The descriptor's box
        // stays the system's default box vectors (context.rs:93).
        std::vector<double> origin(3 * (size_t)ns[0], 0.0);
        nb_->fit_virtual_box(origin.data(), ns[0]);
        for (int k = 0; k < 9; ++k) default_box_[k] = descs[0].box[k];
    }
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
    // Exceptions in global indices (the bonded kernel evaluates them).
    std::vector<int> e1, e2;
    std::vector<double> eqq, esig, eeps, ec6;
    if (nb_->settings().include_direct)
        for (int r = 0; r < nsys; ++r) {
            const NbSettings& c = nb_->settings(r);
            const int off = nb_->atom_offset(r);
            const std::vector<double> c6 = exception_c6(c);
            ec6.insert(ec6.end(), c6.begin(), c6.end());
            for (size_t k = 0; k < c.exc1.size(); ++k) {
                e1.push_back(off + c.exc1[k]);
                e2.push_back(off + c.exc2[k]);
                eqq.push_back(c.exc_qq[k]);
                esig.push_back(c.exc_sig[k]);
                eeps.push_back(c.exc_eps[k]);
            }
        }
    build_bonded(e1, e2, eqq, esig, eeps, b1, b2, blen, bk, a1, a2, a3, ath, ak);
    if (nb_->ljpme()) {
        std::vector<double> p(bonded_exc_order_.size());
        for (size_t k = 0; k < p.size(); ++k) p[k] = ec6[bonded_exc_order_[k]];
        if (p.empty()) p.push_back(0.0);
        bonded_.ec6.upload(p, s_);
    }
    bond1_h_ = b1;
    bond2_h_ = b2;
    setup_constraints(descs, nsys, inv);
    dt_ = descs[0].step_size;
    temperature_ = descs[0].temperature;
    friction_ = descs[0].friction;
    pos_.alloc(3 * (size_t)n_);
    vel_.alloc(3 * (size_t)n_);
    list_src_.alloc(3 * (size_t)n_);
    touched_.assign(nsys, 1);
    fill_async<double>(list_src_.p, 3 * (size_t)n_, 0.0, s_);
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
    snap_list_.alloc(3 * (size_t)n_);
    snap_step_.alloc(1);
    snap_sys_.alloc(nsys);
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
    const size_t off3 = 3 * (size_t)nb_->atom_offset(r), bytes = 3 * (size_t)n * sizeof(double);
    OMM_HIP(hipMemcpyAsync(pos_.p + off3, pos, bytes, hipMemcpyHostToDevice, s_));
    OMM_HIP(hipMemcpyAsync(list_src_.p + off3, pos, bytes, hipMemcpyHostToDevice, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    touched_[r] = 1;
    positions_dirty_ = true;
}

// Every system's list is a function of the positions it was built from and
// the box.  A system set since the last preparation is built from its source
// positions (those set, or the list positions of a restored state); the
// others are rebuilt from the positions of their current lists, which
// reproduces those lists.  The rebuild requests are then those the
// integrators would raise at the current positions.
void MdContext::require_periodic_for_barostat() const {
    // context.rs:93-101
    if (nb_->is_virtual()) throw Error(OMM_ERR_INVALID, "a barostat cannot be used with a non-periodic system");
}

bool MdContext::refit_virtual_box(const double* extra_dev) {
    const size_t n3 = 3 * (size_t)n_;
    std::vector<double> h(extra_dev ? 2 * n3 : n3);
    OMM_HIP(hipMemcpyAsync(h.data(), pos_.p, n3 * sizeof(double), hipMemcpyDeviceToHost, s_));
    if (extra_dev) OMM_HIP(hipMemcpyAsync(h.data() + n3, extra_dev, n3 * sizeof(double), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    if (!nb_->fit_virtual_box(h.data(), (int)(h.size() / 3))) return false;
    drop_graph();  // the device was set up afresh
    return true;
}

void MdContext::prepare() {
    if (!positions_dirty_) return;
    TraceRange range("omm.prepare");
    if (nb_->is_virtual() && refit_virtual_box(nullptr)) {
        // New lists, built from the current positions.
        OMM_HIP(hipMemcpyAsync(list_src_.p, pos_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        std::fill(touched_.begin(), touched_.end(), 1);
    }
    const double* built = nb_->pos_build();
    for (int r = 0; r < nsys_; ++r)
        if (!touched_[r] && built) {
            const size_t off3 = 3 * (size_t)nb_->atom_offset(r);
            OMM_HIP(hipMemcpyAsync(list_src_.p + off3, built + off3, 3 * (size_t)nb_->num_particles(r) * sizeof(double),
                                   hipMemcpyDeviceToDevice, s_));
        }
    if (nb_->calibrate(list_src_.p, status_.p, s_)) drop_graph();
    nb_->recheck_list(pos_.p, s_);
    OMM_HIP(hipStreamSynchronize(s_));
    std::fill(touched_.begin(), touched_.end(), 0);
    positions_dirty_ = false;
}

void MdContext::set_state(int r, int64_t step, const double* box, const double* pos, const double* vel,
                          const double* list_pos) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    if (step < 0) throw Error(OMM_ERR_INVALID, "step count must be nonnegative");
    if (list_pos && !pos) throw Error(OMM_ERR_INVALID, "list positions are restored together with positions");
    if (box && nb_->is_virtual()) throw Error(OMM_ERR_INVALID, "a non-periodic system takes no box vectors");
    const int n = nb_->num_particles(r);
    for (const double* a : {pos, vel, list_pos})
        if (a)
            for (int i = 0; i < 3 * n; ++i)
                if (!std::isfinite(a[i])) throw Error(OMM_ERR_NONFINITE, "state arrays must be finite");
    OMM_HIP(hipStreamSynchronize(s_));
    if (box) nb_->set_box(r, box, s_);
    const size_t off3 = 3 * (size_t)nb_->atom_offset(r), bytes = 3 * (size_t)n * sizeof(double);
    if (pos) {
        OMM_HIP(hipMemcpyAsync(pos_.p + off3, pos, bytes, hipMemcpyHostToDevice, s_));
        OMM_HIP(hipMemcpyAsync(list_src_.p + off3, list_pos ? list_pos : pos, bytes, hipMemcpyHostToDevice, s_));
    } else if (box) {
        // A new box with the current positions: the list follows the box.
        OMM_HIP(hipMemcpyAsync(list_src_.p + off3, pos_.p + off3, bytes, hipMemcpyDeviceToDevice, s_));
    }
    if (pos || box) {
        touched_[r] = 1;
        positions_dirty_ = true;
    }
    if (vel) OMM_HIP(hipMemcpyAsync(vel_.p + off3, vel, bytes, hipMemcpyHostToDevice, s_));
    step_count_ = step;
    fill_async<unsigned long long>(step_dev_.p, 1, (unsigned long long)step, s_);
    OMM_HIP(hipStreamSynchronize(s_));
}

void MdContext::set_capacity(int list_cap, int bin_cap) {
    prepare();
    drop_graph();
    const double* built = nb_->pos_build();
    if (built && nb_->settings().include_direct)
        OMM_HIP(hipMemcpyAsync(list_src_.p, built, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    nb_->set_capacity(list_cap, bin_cap, list_src_.p, status_.p, s_);
}

void MdContext::capacity(int* list_cap, int* bin_cap, int64_t* replays) {
    nb_->capacity(list_cap, bin_cap);
    *replays = overflow_replays_;
}

void MdContext::list_positions(int r, double* out) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    prepare();
    const size_t off3 = 3 * (size_t)nb_->atom_offset(r), bytes = 3 * (size_t)nb_->num_particles(r) * sizeof(double);
    const double* built = nb_->settings().include_direct ? nb_->pos_build() : nullptr;
    OMM_HIP(hipMemcpyAsync(out, (built ? built : pos_.p) + off3, bytes, hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
}

void MdContext::set_tuning(const NbOperator::Tuning& t) {
    OMM_HIP(hipStreamSynchronize(s_));
    prepare();
    nb_->apply_tuning(t, pos_.p, status_.p, s_);
    drop_graph();
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

// Groups the bonded terms by small molecule (connected components of at most
// four atoms over bonds, angles and exceptions); terms of larger components
// stay single.  Groups are ordered by their lowest atom, terms keep their
// order within a group, so the layout is deterministic.
void MdContext::build_bonded(const std::vector<int>& e1, const std::vector<int>& e2, const std::vector<double>& eqq,
                             const std::vector<double>& esig, const std::vector<double>& eeps,
                             const std::vector<int>& b1, const std::vector<int>& b2, const std::vector<double>& blen,
                             const std::vector<double>& bk, const std::vector<int>& a1, const std::vector<int>& a2,
                             const std::vector<int>& a3, const std::vector<double>& ath,
                             const std::vector<double>& ak) {
    std::vector<int> parent(n_), size(n_, 1);
    for (int i = 0; i < n_; ++i) parent[i] = i;
    std::function<int(int)> root = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    auto join = [&](int a, int b) {
        a = root(a);
        b = root(b);
        if (a == b) return;
        if (a > b) std::swap(a, b);
        parent[b] = a;
        size[a] += size[b];
    };
    for (size_t k = 0; k < e1.size(); ++k) join(e1[k], e2[k]);
    for (size_t k = 0; k < b1.size(); ++k) join(b1[k], b2[k]);
    for (size_t k = 0; k < a1.size(); ++k) {
        join(a1[k], a2[k]);
        join(a2[k], a3[k]);
    }
    // Group index of every small component that carries terms.
    std::vector<int> group_of(n_, -1), roots;
    auto small = [&](int a) { return size[root(a)] <= 4; };
    auto mark = [&](int a) {
        const int r = root(a);
        if (group_of[r] < 0) {
            group_of[r] = 0;
            roots.push_back(r);
        }
    };
    for (size_t k = 0; k < e1.size(); ++k)
        if (small(e1[k])) mark(e1[k]);
    for (size_t k = 0; k < b1.size(); ++k)
        if (small(b1[k])) mark(b1[k]);
    for (size_t k = 0; k < a1.size(); ++k)
        if (small(a1[k])) mark(a1[k]);
    std::sort(roots.begin(), roots.end());  // roots are the lowest atoms of their components
    const int ng = (int)roots.size();
    for (int g = 0; g < ng; ++g) group_of[roots[g]] = g;
    std::vector<int> atoms(4 * (size_t)ng, -1), fill(ng, 0);
    for (int i = 0; i < n_; ++i) {
        const int r = root(i);
        if (size[r] <= 4 && group_of[r] >= 0) {
            const int g = group_of[r];
            atoms[4 * g + fill[g]++] = i;
        }
    }
    auto gid = [&](int a) { return small(a) ? group_of[root(a)] : -1; };
    // Stable order: grouped terms by group, then the others.
    auto order = [&](size_t count, auto&& group_fn, std::vector<int>& offsets) {
        std::vector<int> idx(count);
        for (size_t k = 0; k < count; ++k) idx[k] = (int)k;
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            int gx = group_fn(x), gy = group_fn(y);
            gx = gx < 0 ? INT_MAX : gx;
            gy = gy < 0 ? INT_MAX : gy;
            return gx < gy;
        });
        offsets.assign(ng + 1, 0);
        int grouped = 0;
        for (size_t k = 0; k < count; ++k) {
            const int g = group_fn((int)k);
            if (g >= 0) {
                ++offsets[g + 1];
                ++grouped;
            }
        }
        for (int g = 0; g < ng; ++g) offsets[g + 1] += offsets[g];
        return std::make_pair(idx, grouped);
    };
    std::vector<int> oe, ob, oa;
    auto [ie, ge] = order(e1.size(), [&](int k) { return gid(e1[k]); }, oe);
    auto [ib, gb] = order(b1.size(), [&](int k) { return gid(b1[k]); }, ob);
    auto [ia, ga] = order(a1.size(), [&](int k) { return gid(a1[k]); }, oa);
    auto permute = [](const auto& v, const std::vector<int>& idx) {
        std::remove_const_t<std::remove_reference_t<decltype(v)>> out(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) out[k] = v[idx[k]];
        return out;
    };
    BondedDev& B = bonded_;
    B.ngroups = ng;
    B.ne = (int)e1.size();
    B.nb = (int)b1.size();
    B.na = (int)a1.size();
    B.ge = ge;
    B.gb = gb;
    B.ga = ga;
    B.grp_atoms.upload(atoms.empty() ? std::vector<int>{-1, -1, -1, -1} : atoms, s_);
    B.grp_e.upload(oe, s_);
    B.grp_b.upload(ob, s_);
    B.grp_a.upload(oa, s_);
    auto up_i = [&](DevBuf<int>& d, const std::vector<int>& v, const std::vector<int>& idx) {
        std::vector<int> p = permute(v, idx);
        if (p.empty()) p.push_back(0);
        d.upload(p, s_);
    };
    auto up_d = [&](DevBuf<double>& d, const std::vector<double>& v, const std::vector<int>& idx) {
        std::vector<double> p = permute(v, idx);
        if (p.empty()) p.push_back(0.0);
        d.upload(p, s_);
    };
    bonded_exc_order_ = ie;
    up_i(B.e1, e1, ie);
    up_i(B.e2, e2, ie);
    up_d(B.eqq, eqq, ie);
    up_d(B.esig, esig, ie);
    up_d(B.eeps, eeps, ie);
    up_i(B.b1, b1, ib);
    up_i(B.b2, b2, ib);
    up_d(B.blen, blen, ib);
    up_d(B.bk, bk, ib);
    up_i(B.a1, a1, ia);
    up_i(B.a2, a2, ia);
    up_i(B.a3, a3, ia);
    up_d(B.ath, ath, ia);
    up_d(B.ak, ak, ia);
}

void MdContext::launch_bonded(hipStream_t s, long long* force_fx, double* force_d, long long* energy_fx,
                              double* energy_d) {
    const BondedDev& B = bonded_;
    const int total = B.ngroups * GROUP_LANES + (B.ne - B.ge) + (B.nb - B.gb) + (B.na - B.ga);
    if (total == 0) return;
    const int blocks = (total + 255) / 256;
    auto go = [&](auto* force, auto* energy) {
        using A = std::remove_pointer_t<decltype(force)>;
        hipLaunchKernelGGL(k_bonded<A>, dim3(blocks), dim3(256), 0, s, B.ngroups, B.grp_atoms.p, B.grp_e.p,
                           B.grp_b.p, B.grp_a.p, B.ne, B.ge, B.e1.p, B.e2.p, B.eqq.p, B.esig.p, B.eeps.p, B.nb, B.gb,
                           B.b1.p, B.b2.p, B.blen.p, B.bk.p, B.na, B.ga, B.a1.p, B.a2.p, B.a3.p, B.ath.p, B.ak.p,
                           nb_->charge_dev(), nb_->exceptions_pbc() ? 1 : 0, nb_->screened() ? 1 : 0, nb_->alpha(),
                           nb_->ljpme() ? B.ec6.p : nullptr, nb_->ljpme() ? nb_->lj_alpha() : 0.0, pos_.p, nb_->atom_sys_dev(), nb_->systems_dev(), force, energy, status_.p);
    };
    if (mixed_) go(force_fx, energy_fx);
    else go(force_d, energy_d);
    OMM_HIP(hipGetLastError());
}

void MdContext::enqueue_forces(hipStream_t s, bool energy, int rebuild) {
    enqueue_forces_into(s, energy, rebuild, force_fx_.p, force_d_.p, energy_fx_.p, energy_d_.p);
}

void MdContext::enqueue_forces_into(hipStream_t s, bool energy, int rebuild, long long* force_fx, double* force_d,
                                    long long* energy_fx, double* energy_d) {
    const NbSettings& c = nb_->settings();
    nb_->enqueue(pos_.p, c.include_direct, true, energy, rebuild, false, force_fx, force_d, energy_fx, energy_d,
                 status_.p, s, false);
    // Concurrent with the mesh; force-only steps sum no bonded energies.
    launch_bonded(s, force_fx, force_d, energy ? energy_fx : nullptr, energy ? energy_d : nullptr);
    nb_->join(s);
}

// One barostat attempt for every replica (context/program.rs:14-41): the
// energy before the move, the proposal and rigid scaling, the energy after
// (FP64 energy path, list rebuilt for the new boxes), the Metropolis test,
// and the restore of rejected replicas.  The list is rebuilt afterwards.
void MdContext::enqueue_barostat(hipStream_t s) {
    BaroDev& B = baro_;
    const size_t n3 = 3 * (size_t)n_, es = (size_t)E_SLOTS * nsys_;
    const int sb = (nsys_ + 63) / 64;
    if (mixed_) {
        fill_async<long long>(B.force_fx.p, n3, 0ll, s);
        fill_async<long long>(B.e_old_fx.p, es, 0ll, s);
        fill_async<long long>(B.e_new_fx.p, es, 0ll, s);
    } else {
        fill_async<double>(B.force_d.p, n3, 0.0, s);
        fill_async<double>(B.e_old_d.p, es, 0.0, s);
        fill_async<double>(B.e_new_d.p, es, 0.0, s);
    }
    enqueue_forces_into(s, true, NbOperator::REBUILD_FLAG, B.force_fx.p, B.force_d.p, B.e_old_fx.p, B.e_old_d.p);
    hipLaunchKernelGGL(k_baro_propose, dim3(sb), dim3(64), 0, s, nsys_, nb_->systems_dev(), B.st.p, B.saved_box.p,
                       B.scale.p, B.seed, nb_->settings().cutoff, B.kind, B.scale_axes, B.xy_mode, B.z_mode, status_.p);
    hipLaunchKernelGGL(k_baro_scale, dim3((B.nmol + 255) / 256), dim3(256), 0, s, B.nmol, B.mol_start.p,
                       B.mol_atoms.p, B.mol_sys.p, B.scale.p, pos_.p, B.pos_saved.p);
    nb_->force_rebuild(s);
    enqueue_forces_into(s, true, NbOperator::REBUILD_FLAG, B.force_fx.p, B.force_d.p, B.e_new_fx.p, B.e_new_d.p);
    const double kt = BOLTZMANN * B.temperature;
    if (mixed_)
        hipLaunchKernelGGL(k_baro_accept<long long>, dim3(sb), dim3(64), 0, s, nsys_, nb_->systems_dev(), B.st.p,
                           B.saved_box.p, B.rejected.p, B.e_old_fx.p, B.e_new_fx.p, B.nmol_sys.p, B.pressure, B.kind,
                           B.axis_pressure[0], B.axis_pressure[1], B.axis_pressure[2], B.surface_tension, kt);
    else
        hipLaunchKernelGGL(k_baro_accept<double>, dim3(sb), dim3(64), 0, s, nsys_, nb_->systems_dev(), B.st.p,
                           B.saved_box.p, B.rejected.p, B.e_old_d.p, B.e_new_d.p, B.nmol_sys.p, B.pressure, B.kind,
                           B.axis_pressure[0], B.axis_pressure[1], B.axis_pressure[2], B.surface_tension, kt);
    hipLaunchKernelGGL(k_baro_restore, dim3((n_ + 255) / 256), dim3(256), 0, s, n_, nb_->atom_sys_dev(),
                       B.rejected.p, pos_.p, B.pos_saved.p);
    nb_->force_rebuild(s);
    OMM_HIP(hipGetLastError());
}

void MdContext::set_barostat(double pressure, double temperature, int frequency, uint64_t seed) {
    require_periodic_for_barostat();
    // barostat.rs:65-86, 205-218
    if (!std::isfinite(pressure)) throw Error(OMM_ERR_INVALID, "pressure must be finite");
    if (!(std::isfinite(temperature) && temperature >= 0.0))
        throw Error(OMM_ERR_INVALID, "temperature must be nonnegative and finite");
    if (frequency < 0) throw Error(OMM_ERR_INVALID, "barostat frequency must be nonnegative");
    OMM_HIP(hipStreamSynchronize(s_));
    drop_graph();
    BaroDev& B = baro_;
    B.enabled = frequency > 0;
    B.frequency = frequency;
    B.phase = 0;
    B.pressure = pressure;
    B.temperature = temperature;
    B.seed = seed;
    B.kind = 0;
    B.scale_axes = 7;
    // At frequency 0 the barostat makes no attempts but stays set, as a
    // barostat force whose frequency is 0 does, for the pressure queries.
    B.present = true;
    // Molecules: connected components over the bonds (system.rs:100-141).
    std::vector<int> parent(n_);
    for (int i = 0; i < n_; ++i) parent[i] = i;
    auto root = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (size_t k = 0; k < cons1_h_.size(); ++k) {
        const int a = root(cons1_h_[k]), b = root(cons2_h_[k]);
        parent[std::max(a, b)] = std::min(a, b);
    }
    for (size_t k = 0; k < bond1_h_.size(); ++k) {
        const int a = root(bond1_h_[k]), b = root(bond2_h_[k]);
        parent[std::max(a, b)] = std::min(a, b);
    }
    std::vector<std::vector<int>> groups(n_);
    for (int i = 0; i < n_; ++i) groups[root(i)].push_back(i);
    std::vector<int> start{0}, atoms, msys, per_sys(nsys_, 0);
    std::vector<SysInfo> sys(nsys_);
    OMM_HIP(hipMemcpyAsync(sys.data(), nb_->systems_dev(), nsys_ * sizeof(SysInfo), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    std::vector<int> atom_sys(n_);
    for (int r = 0; r < nsys_; ++r)
        for (int i = 0; i < sys[r].n; ++i) atom_sys[sys[r].atom_off + i] = r;
    for (int i = 0; i < n_; ++i) {
        if (groups[i].empty()) continue;
        atoms.insert(atoms.end(), groups[i].begin(), groups[i].end());
        start.push_back((int)atoms.size());
        msys.push_back(atom_sys[i]);
        ++per_sys[atom_sys[i]];
    }
    B.nmol = (int)msys.size();
    B.mol_start.upload(start, s_);
    B.mol_atoms.upload(atoms, s_);
    B.mol_sys.upload(msys, s_);
    B.mol_start_h = start;
    B.mol_atoms_h = atoms;
    B.mol_sys_h = msys;
    B.nmol_sys.upload(per_sys, s_);
    std::vector<BaroState> st(nsys_);
    for (int r = 0; r < nsys_; ++r) {
        st[r] = BaroState{};
        st[r].volume_scale = 0.01 * sys[r].box.v[0] * sys[r].box.v[4] * sys[r].box.v[8];
        for (int a = 0; a < 3; ++a) st[r].axis_volume_scale[a] = st[r].volume_scale;  // anisotropic_barostat.rs:221
    }
    B.st.upload(st, s_);
    B.snap_st.alloc(nsys_);
    B.saved_box.alloc(nsys_);
    B.rejected.alloc(nsys_);
    B.scale.alloc(3 * (size_t)nsys_);
    B.pos_saved.alloc(3 * (size_t)n_);
    const size_t es = (size_t)E_SLOTS * nsys_;
    if (mixed_) {
        B.force_fx.alloc(3 * (size_t)n_);
        B.e_old_fx.alloc(es);
        B.e_new_fx.alloc(es);
    } else {
        B.force_d.alloc(3 * (size_t)n_);
        B.e_old_d.alloc(es);
        B.e_new_d.alloc(es);
    }
}

void MdContext::set_anisotropic_barostat(const double pressure[3], double temperature, int scale_axes, int frequency,
                                         uint64_t seed) {
    require_periodic_for_barostat();
    // anisotropic_barostat.rs:42-74, 217-230
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(pressure[a])) throw Error(OMM_ERR_INVALID, "pressure must be finite");
    if ((scale_axes & 7) == 0 && frequency > 0) throw Error(OMM_ERR_INVALID, "no axis is enabled for scaling");
    set_barostat(pressure[0], temperature, frequency, seed);
    BaroDev& B = baro_;
    B.kind = 1;
    B.scale_axes = scale_axes & 7;
    for (int a = 0; a < 3; ++a) B.axis_pressure[a] = pressure[a];
}

void MdContext::set_flexible_barostat(double pressure, double temperature, int frequency, uint64_t seed) {
    require_periodic_for_barostat();
    // Its trial boxes are triclinic, which the Ewald sum does not take
    // (nonbonded.rs:1087-1097).
    if (nb_->settings().method == OMM_NB_EWALD)
        throw Error(OMM_ERR_INVALID, "Ewald requires a rectangular periodic box; use PME for triclinic boxes");
    set_barostat(pressure, temperature, frequency, seed);
    BaroDev& B = baro_;
    B.kind = 3;
    // flexible_barostat.rs:224-229: length step 0.01 V^(1/3).
    std::vector<BaroState> st(nsys_);
    OMM_HIP(hipMemcpyAsync(st.data(), B.st.p, nsys_ * sizeof(BaroState), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    for (BaroState& b : st) b.volume_scale = 0.01 * std::cbrt(b.volume_scale / 0.01);
    B.st.upload(st, s_);
}

void MdContext::set_membrane_barostat(double pressure, double surface_tension, double temperature, int xy_mode,
                                      int z_mode, int frequency, uint64_t seed) {
    require_periodic_for_barostat();
    // membrane_barostat.rs:61-103, 278-292
    if (!std::isfinite(surface_tension)) throw Error(OMM_ERR_INVALID, "surface tension must be finite");
    if (xy_mode < 0 || xy_mode > 1) throw Error(OMM_ERR_INVALID, "xy mode must be 0 (isotropic) or 1 (anisotropic)");
    if (z_mode < 0 || z_mode > 2)
        throw Error(OMM_ERR_INVALID, "z mode must be 0 (free), 1 (fixed) or 2 (constant volume)");
    set_barostat(pressure, temperature, frequency, seed);
    BaroDev& B = baro_;
    B.kind = 2;
    B.xy_mode = xy_mode;
    B.z_mode = z_mode;
    B.surface_tension = surface_tension;
}

bool MdContext::set_parameter(const std::string& name, double value) {
    OMM_HIP(hipStreamSynchronize(s_));
    bool changed = false;
    const bool found = nb_->set_global_parameter(name, value, &changed);
    if (changed && nb_->settings().include_direct) {
        // The bonded kernel keeps its own (grouped) copy of the exception parameters.
        std::vector<double> qq, sig, eps, c6;
        for (int r = 0; r < nsys_; ++r) {
            const NbSettings& c = nb_->settings(r);
            const std::vector<double> v = exception_c6(c);
            c6.insert(c6.end(), v.begin(), v.end());
            qq.insert(qq.end(), c.exc_qq.begin(), c.exc_qq.end());
            sig.insert(sig.end(), c.exc_sig.begin(), c.exc_sig.end());
            eps.insert(eps.end(), c.exc_eps.begin(), c.exc_eps.end());
        }
        auto permuted = [&](const std::vector<double>& v) {
            std::vector<double> out(bonded_exc_order_.size());
            for (size_t k = 0; k < out.size(); ++k) out[k] = v[bonded_exc_order_[k]];
            if (out.empty()) out.push_back(0.0);
            return out;
        };
        bonded_.eqq.upload(permuted(qq), s_);
        bonded_.esig.upload(permuted(sig), s_);
        bonded_.eeps.upload(permuted(eps), s_);
        if (nb_->ljpme()) bonded_.ec6.upload(permuted(c6), s_);
    }
    return found;
}

void MdContext::barostat_attempt() {
    if (!baro_.enabled) throw Error(OMM_ERR_STATE, "no barostat is set");
    prepare();
    enqueue_barostat(s_);
    check_status();
}

BaroState MdContext::barostat_state(int r) {
    if (r < 0 || r >= nsys_) throw Error(OMM_ERR_INVALID, "replica index out of range");
    if (!baro_.present) throw Error(OMM_ERR_STATE, "no barostat is set");
    BaroState st;
    OMM_HIP(hipMemcpyAsync(&st, baro_.st.p + r, sizeof(BaroState), hipMemcpyDeviceToHost, s_));
    OMM_HIP(hipStreamSynchronize(s_));
    return st;
}

namespace {
constexpr double BAR_NM3_TO_KJ_PER_MOL = 0.0602214076;  // barostat.rs
// (row, column) of the six tensor components xx yy zz xy xz yz (pressure.rs:10).
constexpr int COMPONENTS[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
}  // namespace

void MdContext::require_barostat(int kind, const char* what) const {
    if (!baro_.present || baro_.kind != kind)
        throw Error(OMM_ERR_STATE, std::string("the context has no ") + what);
}

void MdContext::energies_at(const std::vector<double>& transforms, const std::vector<double>& boxes,
                            double* energies) {
    BaroDev& B = baro_;
    prepare();
    const bool ewald = nb_->settings().method == OMM_NB_EWALD;
    for (int r = 0; r < nsys_; ++r) {
        const double* b = boxes.data() + 9 * (size_t)r;
        nb_->check_box(b);
        if (ewald && !(b[3] == 0.0 && b[6] == 0.0 && b[7] == 0.0))
            throw Error(OMM_ERR_INVALID, "Ewald requires a rectangular periodic box; use PME for triclinic boxes");
    }
    B.transform.upload(transforms, s_);
    B.trial_box.upload(boxes, s_);
    const int sb = (nsys_ + 63) / 64;
    OMM_HIP(hipMemcpyAsync(B.pos_saved.p, pos_.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    hipLaunchKernelGGL(k_swap_boxes, dim3(sb), dim3(64), 0, s_, nsys_, nb_->systems_dev(), B.saved_box.p,
                       (const double*)B.trial_box.p);
    if (B.nmol > 0)
        hipLaunchKernelGGL(k_transform_molecules, dim3((B.nmol + 255) / 256), dim3(256), 0, s_, B.nmol,
                           B.mol_start.p, B.mol_atoms.p, B.mol_sys.p, (const double*)B.transform.p, pos_.p);
    OMM_HIP(hipGetLastError());
    nb_->force_rebuild(s_);
    auto restore = [&]() {
        OMM_HIP(hipMemcpyAsync(pos_.p, B.pos_saved.p, 3 * (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        hipLaunchKernelGGL(k_swap_boxes, dim3(sb), dim3(64), 0, s_, nsys_, nb_->systems_dev(), B.saved_box.p,
                           (const double*)nullptr);
        OMM_HIP(hipGetLastError());
        nb_->force_rebuild(s_);
        OMM_HIP(hipStreamSynchronize(s_));
    };
    try {
        evaluate(nullptr, energies, true);
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

void MdContext::kinetic_tensor(std::vector<double>& out) {
    const BaroDev& B = baro_;
    std::vector<double> vel(3 * (size_t)n_);
    vel_.download(vel.data(), vel.size(), s_);
    OMM_HIP(hipStreamSynchronize(s_));
    out.assign(6 * (size_t)nsys_, 0.0);
    for (size_t m = 0; m + 1 < B.mol_start_h.size(); ++m) {
        double p[3] = {0.0, 0.0, 0.0}, mass = 0.0;
        for (int k = B.mol_start_h[m]; k < B.mol_start_h[m + 1]; ++k) {
            const int a = B.mol_atoms_h[k];
            mass += masses_[a];
            for (int d = 0; d < 3; ++d) p[d] += masses_[a] * vel[3 * (size_t)a + d];
        }
        if (mass > 0.0)
            for (int c = 0; c < 6; ++c)
                out[6 * (size_t)B.mol_sys_h[m] + c] += 0.5 * p[COMPONENTS[c][0]] * p[COMPONENTS[c][1]] / mass;
    }
}

void MdContext::current_pressure(double* out) {
    TraceRange range("omm.pressure");
    // pressure.rs:14-42
    require_barostat(0, "MonteCarloBarostat");
    std::vector<double> box(9 * (size_t)nsys_);
    for (int r = 0; r < nsys_; ++r) get_box(r, box.data() + 9 * (size_t)r);
    std::vector<double> T(9 * (size_t)nsys_), trial(box.size()), ep(nsys_), em(nsys_);
    auto at = [&](double scale, double* e) {
        for (int r = 0; r < nsys_; ++r)
            for (int k = 0; k < 9; ++k) {
                T[9 * r + k] = k % 4 == 0 ? scale : 0.0;
                trial[9 * r + k] = box[9 * r + k] * scale;
            }
        energies_at(T, trial, e);
    };
    at(1.001, ep.data());
    at(0.999, em.data());
    std::vector<double> K;
    kinetic_tensor(K);
    for (int r = 0; r < nsys_; ++r) {
        const double* b = box.data() + 9 * (size_t)r;
        const double v = b[0] * b[4] * b[8];
        const double kin = (K[6 * r] + K[6 * r + 1]) + K[6 * r + 2];
        const double dv = v * (1.001 * 1.001 * 1.001 - 0.999 * 0.999 * 0.999);  // powi(3)
        out[r] = ((2.0 / 3.0) * kin / v - (ep[r] - em[r]) / dv) / BAR_NM3_TO_KJ_PER_MOL;
    }
}

void MdContext::axis_pressures(int kind, double* out) {
    // pressure.rs:44-105
    require_barostat(kind, kind == 1 ? "MonteCarloAnisotropicBarostat" : "MonteCarloMembraneBarostat");
    std::vector<double> box(9 * (size_t)nsys_);
    for (int r = 0; r < nsys_; ++r) get_box(r, box.data() + 9 * (size_t)r);
    std::vector<double> K;
    kinetic_tensor(K);
    std::vector<double> T(9 * (size_t)nsys_), trial(box.size()), ep(nsys_), em(nsys_);
    for (int axis = 0; axis < 3; ++axis) {
        auto at = [&](double factor, double* e) {
            double scale[3] = {1.0, 1.0, 1.0};
            scale[axis] = factor;
            for (int r = 0; r < nsys_; ++r)
                for (int k = 0; k < 9; ++k) {
                    T[9 * r + k] = k % 4 == 0 ? scale[k / 4] : 0.0;
                    trial[9 * r + k] = box[9 * r + k] * scale[k % 3];  // barostat.rs:32-40
                }
            energies_at(T, trial, e);
        };
        at(1.001, ep.data());
        at(0.999, em.data());
        for (int r = 0; r < nsys_; ++r) {
            const double* b = box.data() + 9 * (size_t)r;
            const double v = b[0] * b[4] * b[8];
            out[3 * r + axis] =
                (2.0 * K[6 * r + axis] / v - (ep[r] - em[r]) / (0.002 * v)) / BAR_NM3_TO_KJ_PER_MOL;
        }
    }
}

void MdContext::flexible_pressure(double* out) {
    // pressure.rs:133-170
    require_barostat(3, "MonteCarloFlexibleBarostat");
    std::vector<double> box(9 * (size_t)nsys_);
    for (int r = 0; r < nsys_; ++r) get_box(r, box.data() + 9 * (size_t)r);
    std::vector<double> K;
    kinetic_tensor(K);
    std::vector<double> T(9 * (size_t)nsys_), trial(box.size()), e[2] = {std::vector<double>(nsys_),
                                                                         std::vector<double>(nsys_)};
    for (int c = 0; c < 6; ++c) {
        const int column = COMPONENTS[c][0], row = COMPONENTS[c][1];
        std::vector<char> use(nsys_);
        bool any = false;
        for (int r = 0; r < nsys_; ++r) any |= (use[r] = c < 3 || box[9 * r + 3 * row + column] != 0.0);
        for (int sign = 0; sign < 2; ++sign) {
            std::fill(e[sign].begin(), e[sign].end(), 0.0);
            if (!any) continue;
            const double delta = sign == 0 ? 0.001 : -0.001;
            for (int r = 0; r < nsys_; ++r) {
                double* t = trial.data() + 9 * (size_t)r;
                double* m = T.data() + 9 * (size_t)r;
                for (int k = 0; k < 9; ++k) {
                    t[k] = box[9 * r + k];
                    m[k] = k % 4 == 0 ? 1.0 : 0.0;
                }
                if (!use[r]) continue;  // evaluated at its own box, result unused
                t[3 * row + column] *= 1.0 + delta;
                if (!reduce_box(t)) throw Error(OMM_ERR_INVALID, "invalid pressure trial box");
                if (c < 3) m[4 * c] += delta;
            }
            std::vector<double> got(nsys_);
            energies_at(T, trial, got.data());
            for (int r = 0; r < nsys_; ++r)
                if (use[r]) e[sign][r] = got[r];
        }
        for (int r = 0; r < nsys_; ++r) {
            const double* b = box.data() + 9 * (size_t)r;
            const double v = b[0] * b[4] * b[8];
            out[6 * r + c] = (2.0 * K[6 * r + c] - (e[0][r] - e[1][r]) / 0.002) / (v * BAR_NM3_TO_KJ_PER_MOL);
        }
    }
}

void MdContext::stress_tensor(bool include_kinetic, double* out) {
    // pressure.rs:172-209
    require_barostat(3, "MonteCarloFlexibleBarostat");
    std::vector<double> box(9 * (size_t)nsys_);
    for (int r = 0; r < nsys_; ++r) get_box(r, box.data() + 9 * (size_t)r);
    std::vector<double> K(6 * (size_t)nsys_, 0.0);
    if (include_kinetic) kinetic_tensor(K);
    std::vector<double> T(9 * (size_t)nsys_), trial(box.size()), e[2] = {std::vector<double>(nsys_),
                                                                         std::vector<double>(nsys_)};
    for (int c = 0; c < 6; ++c) {
        const int i = COMPONENTS[c][0], j = COMPONENTS[c][1];
        for (int sign = 0; sign < 2; ++sign) {
            const double delta = sign == 0 ? 0.001 : -0.001;
            for (int r = 0; r < nsys_; ++r) {
                double* t = trial.data() + 9 * (size_t)r;
                double* m = T.data() + 9 * (size_t)r;
                for (int k = 0; k < 9; ++k) {
                    t[k] = box[9 * r + k];
                    m[k] = k % 4 == 0 ? 1.0 : 0.0;
                }
                for (int vec = 0; vec < 3; ++vec) {
                    const double shift = delta * t[3 * vec + j];
                    t[3 * vec + i] += shift;
                }
                if (!reduce_box(t)) throw Error(OMM_ERR_INVALID, "invalid stress trial box");
                m[3 * i + j] += delta;
            }
            energies_at(T, trial, e[sign].data());
        }
        for (int r = 0; r < nsys_; ++r) {
            const double* b = box.data() + 9 * (size_t)r;
            const double v = b[0] * b[4] * b[8];
            out[6 * r + c] = ((e[0][r] - e[1][r]) / 0.002 - 2.0 * K[6 * r + c]) / (v * BAR_NM3_TO_KJ_PER_MOL);
        }
    }
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
void MdContext::enqueue_step(hipStream_t s, int rebuild, bool barostat, int offset, int advance) {
    if (barostat) enqueue_barostat(s);
    enqueue_forces(s, false, rebuild);
    const double decay = std::exp(-friction_ * dt_);
    const double variance = BOLTZMANN * temperature_ * -std::expm1(-2.0 * friction_ * dt_);
    const double lim = 0.5 * nb_->skin();
    const int blocks = (n_ + 255) / 256;
    const int check = nb_->settings().include_direct ? 1 : 0;
    const ConstraintDev& C = cons_;
    const unsigned char* constrained = C.ncomp ? C.constrained.p : nullptr;
    if (mixed_)
        hipLaunchKernelGGL(k_lfmiddle<long long>, dim3(blocks), dim3(256), 0, s, n_, pos_.p, vel_.p, force_fx_.p,
                           inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance, step_dev_.p,
                           offset, nb_->pos_build(), lim * lim, check, nb_->list_flag(), constrained, status_.p);
    else
        hipLaunchKernelGGL(k_lfmiddle<double>, dim3(blocks), dim3(256), 0, s, n_, pos_.p, vel_.p, force_d_.p,
                           inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance, step_dev_.p,
                           offset, nb_->pos_build(), lim * lim, check, nb_->list_flag(), constrained, status_.p);
    if (C.ncomp) {
        const int cb = (C.ncomp + 63) / 64;
        auto go = [&](auto* force) {
            using A = std::remove_pointer_t<decltype(force)>;
            hipLaunchKernelGGL(k_constrained_step<A>, dim3(cb), dim3(64), 0, s, C.ncomp, C.atom_off.p, C.atoms.p,
                               C.cons_off.p, C.c1.p, C.c2.p, C.d2.p, C.w1.p, C.w2.p, C.tolerance, 0, pos_.p, vel_.p,
                               force, inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance,
                               step_dev_.p, offset, nb_->pos_build(), lim * lim, check, nb_->list_flag(), C.prev.p,
                               C.proposed.p, status_.p);
        };
        if (mixed_) go(force_fx_.p);
        else go(force_d_.p);
    }
    if (advance > 0) {
        const int eslots = E_SLOTS * nsys_;
        const int eb = std::min(64, (eslots + 255) / 256);
        if (mixed_)
            hipLaunchKernelGGL(k_step_advance<long long>, dim3(eb), dim3(256), 0, s, step_dev_.p, advance,
                               energy_fx_.p, eslots);
        else
            hipLaunchKernelGGL(k_step_advance<double>, dim3(eb), dim3(256), 0, s, step_dev_.p, advance, energy_d_.p,
                               eslots);
    }
    OMM_HIP(hipGetLastError());
}

// Constraints of every replica in global indices (System::add_constraint,
// constraints.rs:35-75 checks), grouped into connected components.
void MdContext::setup_constraints(const omm_md_desc* descs, int nsys, const std::vector<double>& inv) {
    std::vector<int> c1, c2;
    std::vector<double> dist;
    double tol = 0.0;
    for (int r = 0; r < nsys; ++r) {
        const omm_md_desc* d = descs + r;
        const int off = nb_->atom_offset(r), n = d->num_particles;
        if (d->num_constraints < 0) throw Error(OMM_ERR_INVALID, "constraint count must be nonnegative");
        if (d->num_constraints > 0 && (!d->constraint_particle1 || !d->constraint_particle2 || !d->constraint_distance))
            throw Error(OMM_ERR_INVALID, "constraint arrays must not be null");
        const double t = d->constraint_tolerance == 0.0 ? 1e-5 : d->constraint_tolerance;
        if (!(std::isfinite(t) && t > 0.0 && t < 1.0))
            throw Error(OMM_ERR_INVALID, "constraint tolerance must be between zero and one");
        if (r > 0 && t != tol) throw Error(OMM_ERR_INVALID, "the replicas of a batch must share the constraint tolerance");
        tol = t;
        std::set<std::pair<int, int>> seen;
        for (int k = 0; k < d->num_constraints; ++k) {
            const int i = d->constraint_particle1[k], j = d->constraint_particle2[k];
            const double dd = d->constraint_distance[k];
            if (i < 0 || j < 0 || i >= n || j >= n || i == j)
                throw Error(OMM_ERR_INVALID, "constraint particle indices must be distinct and valid");
            if (!(std::isfinite(dd) && dd > 0.0))
                throw Error(OMM_ERR_INVALID, "constraint distance must be positive and finite");
            if (!seen.insert({std::min(i, j), std::max(i, j)}).second)
                throw Error(OMM_ERR_INVALID, "duplicate constraint between the same particles");
            cons1_h_.push_back(off + i);  // molecules (system.rs:117-119) count every constraint
            cons2_h_.push_back(off + j);
            if ((d->masses[i] == 0.0) != (d->masses[j] == 0.0))
                throw Error(OMM_ERR_INVALID, "a constraint cannot connect a massive and massless particle");
            if (d->masses[i] != 0.0) {  // constraints between massless particles are dropped
                c1.push_back(off + i);
                c2.push_back(off + j);
                dist.push_back(dd);
            }
        }
    }
    ConstraintDev& C = cons_;
    C.tolerance = tol;
    C.ncons = (int)c1.size();
    if (C.ncons == 0) return;
    std::vector<int> parent(n_);
    std::iota(parent.begin(), parent.end(), 0);
    auto root = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (int k = 0; k < C.ncons; ++k) {
        const int a = root(c1[k]), b = root(c2[k]);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
    std::map<int, int> comp_of;  // root -> component
    std::vector<std::vector<int>> cons_of, atoms_of;
    for (int k = 0; k < C.ncons; ++k) {  // constraints keep the reference's order within a component
        const int r = root(c1[k]);
        auto it = comp_of.find(r);
        if (it == comp_of.end()) {
            it = comp_of.emplace(r, (int)cons_of.size()).first;
            cons_of.emplace_back();
            atoms_of.emplace_back();
        }
        cons_of[it->second].push_back(k);
    }
    std::vector<unsigned char> constrained(n_, 0);
    for (int a = 0; a < n_; ++a) {
        const auto it = comp_of.find(root(a));
        if (it != comp_of.end()) {
            atoms_of[it->second].push_back(a);
            constrained[a] = 1;
        }
    }
    C.ncomp = (int)cons_of.size();
    std::vector<int> aoff{0}, atoms, coff{0}, p1, p2;
    std::vector<double> d2, w1, w2;
    for (int c = 0; c < C.ncomp; ++c) {
        atoms.insert(atoms.end(), atoms_of[c].begin(), atoms_of[c].end());
        aoff.push_back((int)atoms.size());
        for (int k : cons_of[c]) {
            p1.push_back(c1[k]);
            p2.push_back(c2[k]);
            d2.push_back(dist[k] * dist[k]);
            w1.push_back(inv[c1[k]]);
            w2.push_back(inv[c2[k]]);
        }
        coff.push_back((int)p1.size());
    }
    C.atom_off.upload(aoff, s_);
    C.atoms.upload(atoms, s_);
    C.cons_off.upload(coff, s_);
    C.c1.upload(p1, s_);
    C.c2.upload(p2, s_);
    C.d2.upload(d2, s_);
    C.w1.upload(w1, s_);
    C.w2.upload(w2, s_);
    C.constrained.upload(constrained, s_);
    C.prev.alloc(3 * (size_t)n_);
    C.proposed.alloc(3 * (size_t)n_);
}

void MdContext::project_constraints(double tolerance, bool velocities) {
    // constraints.rs:80-92
    if (!(std::isfinite(tolerance) && tolerance > 0.0 && tolerance < 1.0))
        throw Error(OMM_ERR_INVALID, "constraint tolerance must be finite and between zero and one");
    prepare();
    const ConstraintDev& C = cons_;
    if (!C.ncomp) return;
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    const int cb = (C.ncomp + 63) / 64;
    auto go = [&](auto* force) {
        using A = std::remove_pointer_t<decltype(force)>;
        hipLaunchKernelGGL(k_constrained_step<A>, dim3(cb), dim3(64), 0, s_, C.ncomp, C.atom_off.p, C.atoms.p,
                           C.cons_off.p, C.c1.p, C.c2.p, C.d2.p, C.w1.p, C.w2.p, tolerance, velocities ? 2 : 1,
                           pos_.p, vel_.p, force, inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, 1.0,
                           0.0, step_dev_.p, 0, nb_->pos_build(), 0.0, 0, nb_->list_flag(), C.prev.p, C.proposed.p,
                           status_.p);
    };
    if (mixed_) go(force_fx_.p);
    else go(force_d_.p);
    OMM_HIP(hipGetLastError());
    const unsigned st = read_status();
    if (st & ST_CONSTRAINTS) {
        fill_async<unsigned>(status_.p, 1, 0u, s_);
        OMM_HIP(hipStreamSynchronize(s_));
        throw Error(OMM_ERR_INVALID, "constraint projection is singular or did not converge; check positions");
    }
    if (!velocities) {
        nb_->force_rebuild(s_);  // the positions moved
        OMM_HIP(hipStreamSynchronize(s_));
    }
}

unsigned MdContext::read_status() {
    unsigned st = 0;
    status_.download(&st, 1, s_);
    OMM_HIP(hipStreamSynchronize(s_));
    return st;
}

void MdContext::check_status() {
    const unsigned st = read_status();
    if (st & (ST_LIST_OVERFLOW | ST_BIN_OVERFLOW)) {
        fill_async<unsigned>(status_.p, 1, 0u, s_);
        throw Error(OMM_ERR_OVERFLOW, "neighbour list or PME bin capacity exceeded during the step");
    }
    if (st & ST_NONFINITE) throw Error(OMM_ERR_NONFINITE, "non-finite force or energy on the device");
    if (st & ST_OVERFLOW) throw Error(OMM_ERR_OVERFLOW, "fixed-point force/energy accumulator overflow");
    if (st & ST_CUTOFF_BOX)
        throw Error(OMM_ERR_CUTOFF_BOX, "a barostat move would make the box smaller than twice the cutoff");
    if (st & ST_CONSTRAINTS)
        throw Error(OMM_ERR_INVALID, "constraints failed: a projection was singular or did not converge in 1000 "
                                     "iterations (langevin.rs:215-277)");
    if (st & ST_VBOX) {
        fill_async<unsigned>(status_.p, 1, 0u, s_);
        throw Error(OMM_ERR_OVERFLOW, "atoms of a non-periodic system kept leaving its virtual box");
    }
}

void MdContext::enqueue_snapshot(hipStream_t s) {
    const size_t n3 = 3 * (size_t)n_;
    const double* built = nb_->settings().include_direct ? nb_->pos_build() : nullptr;
    hipLaunchKernelGGL(k_snapshot, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, pos_.p, snap_pos_.p, vel_.p,
                       snap_vel_.p, built, snap_list_.p, n3, step_dev_.p, snap_step_.p);
    OMM_HIP(hipGetLastError());
    OMM_HIP(hipMemcpyAsync(snap_sys_.p, nb_->systems_dev(), nsys_ * sizeof(SysInfo), hipMemcpyDeviceToDevice, s));
    if (baro_.enabled)
        OMM_HIP(hipMemcpyAsync(baro_.snap_st.p, baro_.st.p, nsys_ * sizeof(BaroState), hipMemcpyDeviceToDevice, s));
}

// Puts back the snapshot of enqueue_snapshot and rebuilds every list from the
// positions it was built from at that time, which reproduces it, so the steps
// replayed from here follow the trajectory of a run that never overflowed.
void MdContext::restore_snapshot() {
    const size_t bytes = 3 * (size_t)n_ * sizeof(double);
    // (A virtual box's table is the operator's own: a replay after leaving
    // its region runs in a new box, and nothing else changes the table.)
    if (!nb_->is_virtual())
        OMM_HIP(hipMemcpyAsync(nb_->systems_dev(), snap_sys_.p, nsys_ * sizeof(SysInfo), hipMemcpyDeviceToDevice,
                               s_));
    if (baro_.enabled)
        OMM_HIP(hipMemcpyAsync(baro_.st.p, baro_.snap_st.p, nsys_ * sizeof(BaroState), hipMemcpyDeviceToDevice, s_));
    OMM_HIP(hipMemcpyAsync(pos_.p, snap_pos_.p, bytes, hipMemcpyDeviceToDevice, s_));
    OMM_HIP(hipMemcpyAsync(vel_.p, snap_vel_.p, bytes, hipMemcpyDeviceToDevice, s_));
    OMM_HIP(hipMemcpyAsync(step_dev_.p, snap_step_.p, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s_));
    clear_accumulators(s_);
    fill_async<unsigned>(status_.p, 1, 0u, s_);
    if (nb_->settings().include_direct) {
        OMM_HIP(hipMemcpyAsync(list_src_.p, snap_list_.p, bytes, hipMemcpyDeviceToDevice, s_));
        nb_->calibrate(list_src_.p, status_.p, s_);  // sizes the storage for the rebuilt lists
        nb_->recheck_list(pos_.p, s_);
    }
    OMM_HIP(hipStreamSynchronize(s_));
}

// The captured graph holds a snapshot of the state and graph_length() steps.
// Each step rebuilds the list only when the previous integrator flagged it
// (the build kernels exit at once otherwise).  With a barostat whose period
// fits a graph, the graph spans one period, the attempt opening its last step
// (the phase is aligned before the graph is replayed); longer periods use
// attempt-free graphs of graph_steps_ steps between eager attempt steps.
void MdContext::capture_graph() {
    TraceRange range("omm.capture_graph");
    OMM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_snapshot(s_);
        const int len = graph_length();
        const bool attempt = period_graph();
        for (int k = 0; k < len; ++k)
            enqueue_step(s_, NbOperator::REBUILD_FLAG, attempt && k == len - 1, k + 1, k == len - 1 ? len : 0);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OMM_HIP(hipStreamEndCapture(s_, &graph_));
    OMM_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
}

void MdContext::set_graph_steps(int steps) {
    if (steps < 0 || steps > 1000) throw Error(OMM_ERR_INVALID, "graph step count must be in [0, 1000]");
    if (steps != graph_steps_) drop_graph();
    graph_steps_ = steps;
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

// Runs `count` steps, as one replay of the captured graph (graph: count is
// the graph length) or as eager launches after a snapshot.  A list or PME-bin
// overflow during them restores the snapshot, grows the storage, rebuilds the
// lists the snapshot had and replays the steps eagerly: the trajectory is that
// of a run that never overflowed (the noise is counter-based).
void MdContext::run_steps(int count, bool graph) {
    TraceRange range(graph ? "omm.graph_steps" : "omm.eager_steps");
    BaroDev& B = baro_;
    const int phase0 = B.phase;
    // The attempt opens the step that completes a period (barostat.rs:129-135).
    auto eager = [&]() {
        for (int k = 0; k < count; ++k) {
            enqueue_step(s_, NbOperator::REBUILD_FLAG, B.enabled && B.phase == B.frequency - 1);
            if (B.enabled) B.phase = (B.phase + 1) % B.frequency;
        }
    };
    if (graph) {
        if (!graph_exec_) capture_graph();
        OMM_HIP(hipGraphLaunch(graph_exec_, s_));
        if (B.enabled) B.phase = (B.phase + count) % B.frequency;
    } else {
        enqueue_snapshot(s_);
        eager();
    }
    unsigned st = read_status();
    for (int attempt = 0; st & (ST_LIST_OVERFLOW | ST_BIN_OVERFLOW | ST_VBOX); ++attempt) {
        if (attempt == 8) check_status();  // throws
        TraceRange replay("omm.overflow_replay");
        ++overflow_replays_;
        drop_graph();
        if (st & ST_VBOX) {
            // An atom of a non-periodic system left the virtual box's region:
            // a box for both the snapshot's and the current positions, and
            // the steps replayed in it.
            fill_async<unsigned>(status_.p, 1, 0u, s_);
            refit_virtual_box(snap_pos_.p);  // a new box, or the region re-centred in the same one
        } else {
            nb_->grow(st & ST_BIN_OVERFLOW);  // the lists are sized by the rebuild from the snapshot
        }
        restore_snapshot();
        B.phase = phase0;
        eager();
        st = read_status();
    }
}

void MdContext::step(int nsteps) {
    if (nsteps < 0) throw Error(OMM_ERR_INVALID, "step count must be nonnegative");
    TraceRange range("omm.step");
    prepare();
    const BaroDev& B = baro_;
    const int len = graph_length();
    const bool graphs = mixed_ && graph_steps_ > 0 && len > 0;
    const int chunk = graph_steps_ > 0 ? graph_steps_ : 10;  // eager steps per snapshot
    int done = 0;
    while (done < nsteps) {
        const int left = nsteps - done;
        // A period graph starts at an attempt boundary; an attempt-free graph
        // ends before the next attempt step.
        const bool fits = !B.enabled || (period_graph() ? B.phase == 0 : B.phase + len <= B.frequency - 1);
        int k;
        bool g = false;
        if (graphs && fits && left >= len) {
            k = len;
            g = true;
        } else if (graphs && B.enabled && period_graph()) {
            k = std::min(left, B.frequency - B.phase);  // up to the next period boundary
        } else if (graphs && B.enabled) {
            k = std::min(left, B.phase < B.frequency - 1 ? B.frequency - 1 - B.phase : 1);  // up to and over the attempt
        } else {
            k = std::min(left, graphs ? len : chunk);
        }
        run_steps(k, g);
        done += k;
    }
    step_count_ += nsteps;
    check_status();
}

void MdContext::evaluate(double* forces, double* energy, bool want_energy) {
    TraceRange range("omm.evaluate");
    prepare();
    for (int attempt = 0;; ++attempt) {
        clear_accumulators(s_);
        enqueue_forces(s_, want_energy, NbOperator::REBUILD_FLAG);
        const unsigned st = read_status();
        if (!(st & (ST_LIST_OVERFLOW | ST_BIN_OVERFLOW)) || attempt == 4) break;
        if (attempt == 3)
            throw Error(OMM_ERR_OVERFLOW, "evaluation kept overflowing (status " + std::to_string(st) + ")");
        fill_async<unsigned>(status_.p, 1, 0u, s_);
        nb_->grow(st);
        nb_->force_rebuild(s_);
        drop_graph();
    }
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
    if (nb_->is_virtual()) {
        bool any = false;
        for (int k = 0; k < 9; ++k) any = any || default_box_[k] != 0.0;
        if (!any) throw Error(OMM_ERR_STATE, "this non-periodic system has no box vectors");
        for (int k = 0; k < 9; ++k) box[k] = default_box_[k];
        return;
    }
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
    ms[6] = timed([&] { launch_bonded(s_, force_fx_.p, force_d_.p, nullptr, nullptr); });
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
                               step_dev_.p, 1, nb_->pos_build(), lim * lim, 1, scratch_flag.p, nullptr, status_.p);
        else
            hipLaunchKernelGGL(k_lfmiddle<double>, dim3((n_ + 255) / 256), dim3(256), 0, s_, n_, px.p, vx.p,
                               force_d_.p, inv_mass_.p, nb_->atom_sys_dev(), nb_->systems_dev(), dt_, decay, variance,
                               step_dev_.p, 1, nb_->pos_build(), lim * lim, 1, scratch_flag.p, nullptr, status_.p);
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
