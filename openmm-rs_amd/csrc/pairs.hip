// Direct-space LJ + Ewald-screened Coulomb over the tile list, and the
// exception kernel (nonbonded.rs:842-924, 995-1036).
//
// One wave64 per task = (cluster of 8 i-atoms, up to TPT consecutive tiles
// of 64 entries of the cluster's list); the loads of the next tiles are in
// flight while a tile is evaluated.  Lane l owns i-atom l%8 and the group of entries
// 8*(l/8) .. 8*(l/8)+7; in step k (0..7) it pairs with entry 8*(l/8) + (l%8 ^ k),
// so every entry is touched by exactly one lane per step.  j-data sit in LDS
// (read-only); the j-force accumulators travel with their entry between the
// lanes of a group through DPP permutations (lane ^ 1, ^ 3, ^ 7), so no lane
// writes LDS inside the loop and no LDS instruction moves them.  Positions are FP64 in HBM; each task converts them to
// FP32 offsets from the i-block's FP64 centre, so deltas carry ~1e-7 nm
// absolute error instead of the ~5e-7 nm of absolute FP32 coordinates.
// Near-cutoff pairs (|r^2 - rc^2| < 1e-5 rc^2) are re-decided in FP64 with
// the reference's exact arithmetic, which makes the pair set bit-exact.
//
// Energy mode: forces stay on the FP32 path, and each pair's energy is
// evaluated in FP64 from FP64 offsets.  The Coulomb classes of a water box
// (O-O, O-H, H-H) are each ~1e6 kJ/mol and cancel to ~1e2-1e3, so any
// distance-dependent FP32 rounding bias would survive in the sum; FP64 pair
// energies remove it.  The MD step does not request energies and never pays
// for this.
#include <climits>
#include <type_traits>

#include "fmath.h"
#include "pairs.h"
#include "terms.h"

namespace ommhip {

namespace {

template <class R>
struct alignas(4 * sizeof(R)) Q4 {
    R x, y, z, w;
};
template <class R>
struct alignas(2 * sizeof(R)) Q2 {
    R x, y;
};

// The reference's cutoff decision for the pair (a, b): delta = x_j - x_i with
// i < j, reduced-triclinic minimum image, r^2 = dx dx + dy dy + dz dz, no FMA.
// Inlined: a call would make the compiler wait for every load in flight
// (the next tile's prefetch) before the band loop of each tile.
__device__ __forceinline__ bool within_cutoff_exact(const double* __restrict__ pos, int a, int b, const Box& bx,
                                                 int periodic, double cutoff2) {
#pragma clang fp contract(off)
    int i = min(a, b), j = max(a, b);
    double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
    if (periodic) min_image_exact(d, bx.v);
    double r2 = d[0] * d[0];
    r2 = r2 + d[1] * d[1];
    r2 = r2 + d[2] * d[2];
    return r2 < cutoff2;
}

// The reference's LJ switch (nonbonded.rs:887-899): S(t) = 1 + t^3 (-10 + t (15 - 6 t))
// and dS/dr for t = (r - r_switch) / (r_cut - r_switch), applied past r_switch.
template <class R>
__device__ __forceinline__ void lj_switch(R r, const PairParams& P, R& s, R& ds) {
    const R inv_width = (R)P.inv_switch_width;
    const R t = (r - (R)P.switch_distance) * inv_width;
    s = (R)1 + t * t * t * ((R)-10 + t * ((R)15 - (R)6 * t));
    ds = (R)-30 * t * t * ((R)1 - t) * ((R)1 - t) * inv_width;
}

// FP64 energy of one pair (nonbonded.rs:1203-1214, 1219-1273) from offsets.
template <int KIND>
__device__ __forceinline__ double pair_energy_fp64(double dx, double dy, double dz, double sig, double eps4,
                                                   double ss, double qq, const PairParams& P) {
    constexpr int ELEC = KIND & ELEC_MASK;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    double e = 0.0;
    if (eps4 != 0.0) {
        const double s2 = sig * sig / r2;
        const double s6 = s2 * s2 * s2;
        e = eps4 * (s6 * s6 - s6);
        if (KIND & PAIR_LJPME) {
            // nonbonded.rs:875-886: + the reciprocal dispersion of the pair,
            // minus both terms at the cutoff.  C6 = 4 sqrt(eps_i eps_j)
            // (sigma_i sigma_j)^3 = eps4 (4 ss)^3 with ss = (sigma_i/2)(sigma_j/2).
            const double c6 = eps4 * 64.0 * ss * ss * ss, rc2 = P.cutoff * P.cutoff, a = P.lj_alpha;
            double dd;
            const double sc2 = sig * sig / rc2, sc6 = sc2 * sc2 * sc2;
            e += disp_recip_pair(c6, r2, a, &dd) - (eps4 * (sc6 * sc6 - sc6) + disp_recip_pair(c6, rc2, a, &dd));
        }
        if ((KIND & PAIR_SWITCH) && r > P.switch_distance) {
            double sw, dsw;
            lj_switch<double>(r, P, sw, dsw);
            e *= sw;
        }
    }
    if (qq != 0.0) {
        if (ELEC == ELEC_EWALD) e += qq * erfc(P.alpha * r) / r;
        else if (ELEC == ELEC_REACTION_FIELD) e += qq * (1.0 / r + P.krf * r2 - P.crf);
        else e += qq / r;
    }
    return e;
}

// Lane permutations by DPP inside each row of 16 lanes: quad_perm [1,0,3,2]
// is lane ^ 1, quad_perm [3,2,1,0] is lane ^ 3, row_half_mirror is lane ^ 7
// (each group of 8 reversed), row_ror:8 is lane ^ 8.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR3 = 0x1B, DPP_XOR7 = 0x141, DPP_XOR8 = 0x128;

// Every control used here reads a lane of the same row, so no lane falls
// back to an old value and the move needs none.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The j-force accumulators move from step k to step k + 1: lane l takes the
// accumulator of lane l ^ (k ^ (k + 1)) of its group of 8.
template <int K, class R>
__device__ __forceinline__ void rotate_j(R& gx, R& gy, R& gz) {
    constexpr int m = K ^ (K + 1);
    constexpr int ctrl = m == 1 ? DPP_XOR1 : (m == 3 ? DPP_XOR3 : DPP_XOR7);
    gx = dpp<ctrl>(gx);
    gy = dpp<ctrl>(gy);
    gz = dpp<ctrl>(gz);
}

// Raw data of one tile entry as loaded from HBM, before it is staged in LDS.
template <class R, bool E64>
struct EntryRaw {
    double x = 0, y = 0, z = 0;
    R q = 0, s = 0, e = 0;
    double qd = 0, sd = 0, ed = 0;  // the energy path's FP64 parameters
};

// Loads the data of atom a >= 0 with no branch: the entries past the end of
// a list load some real atom's data and are masked when their tile is staged,
// so the loads of a tile never wait on a branch's merge.
template <class R, bool E64>
__device__ __forceinline__ void load_entry(const int a, const double* __restrict__ pos, const R* __restrict__ qk,
                                           const R* __restrict__ sh, const R* __restrict__ e2,
                                           const double* __restrict__ qk_d, const double* __restrict__ sh_d,
                                           const double* __restrict__ e2_d, EntryRaw<R, E64>& r) {
    r.x = pos[3 * a];
    r.y = pos[3 * a + 1];
    r.z = pos[3 * a + 2];
    r.q = qk[a];
    r.s = sh[a];
    r.e = e2[a];
    if (E64) {
        r.qd = qk_d[a];
        r.sd = sh_d[a];
        r.ed = e2_d[a];
    }
}

template <class R, class A, int KIND, bool STATS, bool ENERGY>
__global__ __launch_bounds__(64) void k_pairs(const double* __restrict__ pos, const R* __restrict__ qk,
                                              const R* __restrict__ sh, const R* __restrict__ e2,
                                              const double* __restrict__ qk_d, const double* __restrict__ sh_d,
                                              const double* __restrict__ e2_d, const int* __restrict__ sorted_atom,
                                              const unsigned char* __restrict__ clist_ex,
                                              const double* __restrict__ center,
                                              const int* __restrict__ minimg, const int2* __restrict__ clist,
                                              const int* __restrict__ ccount, const int* __restrict__ tasks,
                                              const int* __restrict__ ntasks, const int* __restrict__ block_sys,
                                              const SysInfo* __restrict__ sys, PairParams P, A* fslot, A* energy,
                                              unsigned* status, unsigned long long* stats, int* list_flag,
                                              int nflags) {
    using M = Math<R>;
    constexpr int ELEC = KIND & ELEC_MASK;
    constexpr bool SWITCH = (KIND & PAIR_SWITCH) != 0;
    constexpr bool LJPME = (KIND & PAIR_LJPME) != 0;
    constexpr bool E64 = ENERGY && sizeof(R) == 4;  // separate FP64 energy path
    __shared__ Q4<R> s_pq[JTILE];
    __shared__ Q2<R> s_se[JTILE];
    __shared__ __align__(16) int s_js[JTILE];
    __shared__ int s_ja[JTILE];
    __shared__ __align__(8) unsigned char s_ex[JTILE];  // exclusion bits of each entry (bit k: i-atom k)
    __shared__ Q4<double> s_pd[E64 ? JTILE : 1];
    __shared__ Q2<double> s_sd[E64 ? JTILE : 1];

    const int lane = threadIdx.x;
    const int grp = lane / CLUSTER;  // the lane's group of 8 tile entries
    const int il = lane % CLUSTER;   // the lane's i-atom in the cluster
    const double cutoff2 = P.cutoff * P.cutoff;
    const double band = sizeof(R) == 4 ? 1e-5 : 1e-12;
    const R rc2_lo = (R)(cutoff2 * (1.0 - band));
    const R rc2_hi = (R)(cutoff2 * (1.0 + band));
    const R alpha = (R)P.alpha;
    const R two_a_sqrtpi = (R)(2.0 * P.alpha / sqrt(PI));
    const R krf = (R)P.krf, crf = (R)P.crf;
    const R lj_a2 = (R)(P.lj_alpha * P.lj_alpha);
    // This kernel runs after every list build and before the next integrator,
    // so it serves the rebuild requests.
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < nflags; i += blockDim.x) list_flag[i] = 0;

    const int total = ntasks[0];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int task = tasks[t];
        const int C = task >> 8;
        const int first = task & 255;
        const int I = C / CPB;
        const int count = ccount[C];
        const int ntile = min(TPT, (count + JTILE - 1) / JTILE - first);
        const int sidx = block_sys[I];
        const Box& bx = sys[sidx].box;
        const double c0 = center[3 * I], c1 = center[3 * I + 1], c2 = center[3 * I + 2];
        const int2* list = clist + (size_t)C * P.cap + (size_t)first * JTILE;
        const unsigned char* list_ex = clist_ex + (size_t)C * P.cap + (size_t)first * JTILE;
        const int nentries = count - first * JTILE;
        // Entry e of tile tl is live when tl < ntile and e < nentries.  Every
        // lane loads some real entry of the list (the last one past its end)
        // and the liveness is applied when the tile is staged, so no load
        // sits behind a branch.
        const int last = max(nentries - 1, 0);
        auto live = [&](int tl) { return tl < ntile && tl * JTILE + lane < nentries; };
        auto entry = [&](int tl) { return list[min(tl * JTILE + lane, last)]; };
        auto entry_ex = [&](int tl) -> unsigned { return list_ex[min(tl * JTILE + lane, last)]; };
        // The task waits for two rounds of loads before its first tile: the
        // i-atom's slot with the first two tiles' entries, then the i-atom's
        // data with tile 0's coordinates and parameters.
        const int is = C * CLUSTER + il;
        const int ia = sorted_atom[is];  // -1 in the padding of a system's last block
        // Software pipeline: the entries of tile k + 2 and the coordinates of
        // tile k + 1 are in flight while tile k is evaluated.
        int2 e0 = entry(0), e1 = entry(1);
        unsigned x0 = entry_ex(0), x1 = entry_ex(1);
        const int ia0 = max(ia, 0);
        const double pxi = pos[3 * ia0], pyi = pos[3 * ia0 + 1], pzi = pos[3 * ia0 + 2];
        const R qi_raw = qk[ia0], si_raw = sh[ia0], ei_raw = e2[ia0];
        double qid_raw = 0, sid_raw = 0, eid_raw = 0;
        if (E64) {
            qid_raw = qk_d[ia0];
            sid_raw = sh_d[ia0];
            eid_raw = e2_d[ia0];
        }
        EntryRaw<R, E64> r0;
        load_entry<R, E64>(max(e0.y, 0), pos, qk, sh, e2, qk_d, sh_d, e2_d, r0);

        R xi = 0, yi = 0, zi = 0, qi = 0, si = 0, ei = 0;
        double xid = 0, yid = 0, zid = 0, qid = 0, sid = 0, eid = 0;
        if (ia >= 0) {
            double d[3] = {pxi - c0, pyi - c1, pzi - c2};
            if (P.periodic) wrap_nearest(d, bx);
            xi = (R)d[0];
            yi = (R)d[1];
            zi = (R)d[2];
            qi = qi_raw;
            si = si_raw;
            ei = ei_raw;
            if (E64) {
                xid = d[0];
                yid = d[1];
                zid = d[2];
                qid = qid_raw;
                sid = sid_raw;
                eid = eid_raw;
            }
        }
        // The pair steps read the i-atom's parameters inside divergent
        // branches, where the compiler cannot prove an earlier wait and
        // would wait for every load in flight, the next tile's prefetch
        // included, in each step.  Reading them here once, on every path,
        // completes their loads before the tiles start.
        asm volatile("" ::"v"(qi), "v"(si), "v"(ei));
        if (E64) asm volatile("" ::"v"(qid), "v"(sid), "v"(eid));
        const bool mi = P.periodic && minimg[I];
        R bv[9], ib0 = 0, ib4 = 0, ib8 = 0;
        if (mi) {
            for (int k = 0; k < 9; ++k) bv[k] = (R)bx.v[k];
            ib0 = (R)(1.0 / bx.v[0]);
            ib4 = (R)(1.0 / bx.v[4]);
            ib8 = (R)(1.0 / bx.v[8]);
        }

        R fx = 0, fy = 0, fz = 0;
        double eacc = 0.0;
        unsigned long long npair = 0, hsum = 0;
        for (int k = 0; k < ntile; ++k) {
            const int2 e2n = entry(k + 2);
            const unsigned x2n = entry_ex(k + 2);
            EntryRaw<R, E64> r1;
            load_entry<R, E64>(max(e1.y, 0), pos, qk, sh, e2, qk_d, sh_d, e2_d, r1);

            // Stage tile k in LDS as offsets from the block centre.
            {
                if (!live(k)) {
                    e0 = make_int2(INT_MAX, -1);
                    x0 = 0u;
                }
                Q4<R> pq = {0, 0, 0, 0};
                Q2<R> se = {0, 0};
                Q4<double> pd = {0, 0, 0, 0};
                Q2<double> sd = {0, 0};
                if (e0.y >= 0) {
                    double d[3] = {r0.x - c0, r0.y - c1, r0.z - c2};
                    if (P.periodic) wrap_nearest(d, bx);
                    pq = {(R)d[0], (R)d[1], (R)d[2], r0.q};
                    se = {r0.s, r0.e};
                    if (E64) {
                        pd = {d[0], d[1], d[2], r0.qd};
                        sd = {r0.sd, r0.ed};
                    }
                }
                s_pq[lane] = pq;
                s_se[lane] = se;
                s_js[lane] = e0.x;
                s_ja[lane] = e0.y;
                s_ex[lane] = (unsigned char)x0;
                if (E64) {
                    s_pd[lane] = pd;
                    s_sd[lane] = sd;
                }
            }
            __syncthreads();

            // Valid entries of the lane's group: present, not excluded, and
            // after the i-atom where they are of its own cluster (entries
            // before the cluster's first slot come from blocks the cluster
            // takes cyclically).
            unsigned vmask = 0;
            if (ia >= 0) {
                const int4* g4 = reinterpret_cast<const int4*>(s_js + grp * CLUSTER);
                const int4 lo = g4[0], hi = g4[1];
                const int jv[CLUSTER] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
                const int cfirst = is & ~(CLUSTER - 1);
#pragma unroll
                for (int s8 = 0; s8 < CLUSTER; ++s8)
                    if (jv[s8] != INT_MAX && (jv[s8] > is || jv[s8] < cfirst)) vmask |= 1u << s8;
                // Bit il of each of the group's eight exclusion bytes, gathered
                // into bits 0-7: each byte's bit il moves to the byte's bit 0,
                // and a multiply by 2^0 + 2^7 + 2^14 + 2^21 carries the bits of
                // bytes 0-3 (positions 0, 8, 16, 24) to positions 21-24
                // without collisions.
                const unsigned long long xb = *reinterpret_cast<const unsigned long long*>(s_ex + grp * CLUSTER);
                const unsigned long long sel = (xb >> il) & 0x0101010101010101ull;
                const unsigned lo4 = (((unsigned)sel * 0x00204081u) >> 21) & 0xFu;
                const unsigned hi4 = (((unsigned)(sel >> 32) * 0x00204081u) >> 21) & 0xFu;
                vmask &= ~(lo4 | (hi4 << 4));
            }

            R gx = 0, gy = 0, gz = 0;
            // Offset of the entry at pj from the lane's i-atom, with the FP32
            // minimum image when MI (the block is near half the box).
            auto delta = [&](const Q4<R>& pj, R& dx, R& dy, R& dz, auto MI) {
                dx = pj.x - xi;
                dy = pj.y - yi;
                dz = pj.z - zi;
                if (decltype(MI)::value) {
                    R s = M::round(dz * ib8);
                    dx -= s * bv[6];
                    dy -= s * bv[7];
                    dz -= s * bv[8];
                    s = M::round(dy * ib4);
                    dx -= s * bv[3];
                    dy -= s * bv[4];
                    s = M::round(dx * ib0);
                    dx -= s * bv[0];
                }
            };
            // F/r of a pair at squared distance r2 with the entry's charge qj
            // and LJ parameters sj; e receives the pair's working-precision energy.
            auto pair_fe = [&](const R qj, const Q2<R> sj, const R r2, R& e) -> R {
                const R inv_r = M::rsqrt(r2);
                const R inv_r2 = inv_r * inv_r;
                // Lennard-Jones with Lorentz-Berthelot combination (nonbonded.rs:869-874, 1203-1214).
                const R sig = si + sj.x;
                const R eps4 = ei * sj.y;
                const R s2 = sig * sig * inv_r2;
                const R s6 = s2 * s2 * s2;
                const R s12 = s6 * s6;
                e = eps4 * (s12 - s6);
                R f = eps4 * ((R)6 * s6 - (R)12 * s12) * inv_r2;
                if (LJPME) {
                    // nonbonded.rs:875-886 with C6 = eps4 (4 (sigma_i/2)(sigma_j/2))^3.
                    // Closed form (no small-x series: x = alpha^2 r^2 < 0.01 needs
                    // r < 0.1/alpha, far inside any non-excluded pair's LJ core).
                    const R ss = si * sj.x;
                    const R c6 = eps4 * (R)64 * ss * ss * ss;
                    const R x = lj_a2 * r2;
                    const R ex = M::exp_neg(x);
                    const R poly = (R)1 + x * ((R)1 + (R)0.5 * x);
                    const R c6r6 = c6 * inv_r2 * inv_r2 * inv_r2;
                    f -= (R)6 * c6r6 * inv_r2 * ((R)1 - ex * (poly + x * x * x * (R)(1.0 / 6.0)));
                    if (ENERGY && !E64) {
                        // The plain form's energy (the mixed form's is the FP64 pair energy).
                        const R rc2 = (R)(P.cutoff * P.cutoff);
                        const R sc2 = sig * sig / rc2, sc6 = sc2 * sc2 * sc2;
                        const R xc = lj_a2 * rc2;
                        const R shift = eps4 * (sc6 * sc6 - sc6) +
                                        c6 / (rc2 * rc2 * rc2) *
                                            ((R)1 - M::exp_neg(xc) * ((R)1 + xc * ((R)1 + (R)0.5 * xc)));
                        e += c6r6 * ((R)1 - ex * poly) - shift;
                    }
                }
                if (SWITCH) {
                    // Inside the switching distance t = 0 gives S = 1 and
                    // dS/dr = 0 exactly, so no branch is needed.
                    const R r = M::max(r2 * inv_r, (R)P.switch_distance);
                    R sw, dsw;
                    lj_switch<R>(r, P, sw, dsw);
                    f = sw * f + e * dsw * inv_r;
                    e *= sw;
                }
                const R qq = qi * qj;
                if (ELEC == ELEC_EWALD) {
                    // screened_coulomb (nonbonded.rs:1263-1273)
                    const R r = r2 * inv_r;
                    R erfc_v, ex2;
                    M::erfc_exp(alpha * r, erfc_v, ex2);
                    const R ec = qq * erfc_v * inv_r;
                    e += ec;
                    f -= (ec + qq * two_a_sqrtpi * ex2) * inv_r2;
                } else if (ELEC == ELEC_REACTION_FIELD) {
                    // reaction_field (nonbonded.rs:1233-1250)
                    e += qq * (inv_r + krf * r2 - crf);
                    f += qq * ((R)2 * krf - inv_r * inv_r2);
                } else {
                    // coulomb_pair (nonbonded.rs:1219-1226)
                    const R ec = qq * inv_r;
                    e += ec;
                    f -= ec * inv_r2;
                }
                return f;
            };
            // Energy and statistics of an accepted pair (their variants only).
            auto pair_extra = [&](const int slot, const R e) {
                if (E64) {
                    const Q4<double> pjd = s_pd[slot];
                    const Q2<double> sjd = s_sd[slot];
                    double ddx = pjd.x - xid, ddy = pjd.y - yid, ddz = pjd.z - zid;
                    if (mi) {
                        double dd[3] = {ddx, ddy, ddz};
                        wrap_nearest(dd, bx);
                        ddx = dd[0];
                        ddy = dd[1];
                        ddz = dd[2];
                    }
                    eacc += pair_energy_fp64<KIND>(ddx, ddy, ddz, sid + sjd.x, eid * sjd.y, sid * sjd.x, qid * pjd.w,
                                                   P);
                } else if (ENERGY) {
                    eacc += (double)e;
                }
                if (STATS) {
                    const int ja = s_ja[slot];
                    ++npair;
                    hsum += pair_key(min(ia, ja), max(ia, ja));
                }
            };
            // Step k: lane (grp, il) pairs i-atom il with entry grp * 8 + (il ^ k);
            // each entry meets every i-atom of the cluster once over the 8 steps.
            // Pairs whose FP32 r^2 falls in the band around the cutoff are only
            // marked here and decided in FP64 after the steps.  Every lane
            // evaluates its slot whether or not it holds an accepted pair: a
            // wave nearly always holds one in each step, so a branch would skip
            // nothing, and straight-line steps let the compiler issue the next
            // step's LDS reads early.  A rejected slot's r^2 is replaced by 1
            // (away from the singularity at 0) and its force selected to zero.
            // The minimum-image choice is made once per tile, outside the
            // steps, so the eight steps form one straight block.
            unsigned near = 0;
            auto pair_step = [&](const int k, auto MI) {
                const int s8 = il ^ k;
                const int slot = grp * CLUSTER + s8;
                const Q4<R> pj = s_pq[slot];
                const Q2<R> sj = s_se[slot];
                R dx, dy, dz;
                delta(pj, dx, dy, dz, MI);
                const R r2 = dx * dx + dy * dy + dz * dz;
                const bool valid = (vmask >> s8) & 1u;
                const bool in = valid && (!P.use_cutoff || r2 < rc2_lo);
                if (valid && !in && r2 < rc2_hi) near |= 1u << s8;
                R e;
                R f = pair_fe(pj.w, sj, in ? r2 : (R)1, e);
                f = in ? f : (R)0;
                if ((ENERGY || STATS) && in) pair_extra(slot, e);
                fx += f * dx;
                fy += f * dy;
                fz += f * dz;
                gx -= f * dx;
                gy -= f * dy;
                gz -= f * dz;
            };
            auto eight_steps = [&](auto MI) {
                pair_step(0, MI);
                rotate_j<0>(gx, gy, gz);
                pair_step(1, MI);
                rotate_j<1>(gx, gy, gz);
                pair_step(2, MI);
                rotate_j<2>(gx, gy, gz);
                pair_step(3, MI);
                rotate_j<3>(gx, gy, gz);
                pair_step(4, MI);
                rotate_j<4>(gx, gy, gz);
                pair_step(5, MI);
                rotate_j<5>(gx, gy, gz);
                pair_step(6, MI);
                rotate_j<6>(gx, gy, gz);
                pair_step(7, MI);
            };
            if (mi) eight_steps(std::true_type{});
            else eight_steps(std::false_type{});
            // The cutoff band: the reference's FP64 decision (rare pairs; their
            // j-forces go straight to the slot-ordered accumulator).
            while (__ballot(near != 0u)) {
                if (near) {
                    const int s8 = __ffs(near) - 1;
                    near &= near - 1u;
                    const int slot = grp * CLUSTER + s8;
                    if (within_cutoff_exact(pos, ia, s_ja[slot], bx, P.periodic, cutoff2)) {
                        const Q4<R> pj = s_pq[slot];
                        R dx, dy, dz, e;
                        if (mi) delta(pj, dx, dy, dz, std::true_type{});
                        else delta(pj, dx, dy, dz, std::false_type{});
                        const R f = pair_fe(pj.w, s_se[slot], dx * dx + dy * dy + dz * dz, e);
                        if (ENERGY || STATS) pair_extra(slot, e);
                        fx += f * dx;
                        fy += f * dy;
                        fz += f * dz;
                        const int js = s_js[slot];
                        acc_add(fslot + js, (double)(-f * dx), FORCE_SCALE, status);
                        acc_add(fslot + P.nslots + js, (double)(-f * dy), FORCE_SCALE, status);
                        acc_add(fslot + 2 * P.nslots + js, (double)(-f * dz), FORCE_SCALE, status);
                    }
                }
            }
            // After step 7 the lane holds the accumulator of entry grp * 8 + (il ^ 7).
            const int jmine = grp * CLUSTER + (il ^ (CLUSTER - 1));
            if (s_ja[jmine] >= 0) {
                const int js = s_js[jmine];
                acc_add(fslot + js, (double)gx, FORCE_SCALE, status);
                acc_add(fslot + P.nslots + js, (double)gy, FORCE_SCALE, status);
                acc_add(fslot + 2 * P.nslots + js, (double)gz, FORCE_SCALE, status);
            }
            __syncthreads();
            e0 = e1;
            e1 = e2n;
            x0 = x1;
            x1 = x2n;
            r0 = r1;
        }
        // i-forces: the eight groups' partial sums, in a fixed order.
        fx += dpp<DPP_XOR8>(fx);
        fy += dpp<DPP_XOR8>(fy);
        fz += dpp<DPP_XOR8>(fz);
        fx += __shfl_xor(fx, 16, 64);
        fy += __shfl_xor(fy, 16, 64);
        fz += __shfl_xor(fz, 16, 64);
        fx += __shfl_xor(fx, 32, 64);
        fy += __shfl_xor(fy, 32, 64);
        fz += __shfl_xor(fz, 32, 64);
        if (grp == 0 && ia >= 0) {
            acc_add(fslot + is, (double)fx, FORCE_SCALE, status);
            acc_add(fslot + P.nslots + is, (double)fy, FORCE_SCALE, status);
            acc_add(fslot + 2 * P.nslots + is, (double)fz, FORCE_SCALE, status);
        }
        if (ENERGY) {
            double etot = wave_sum(eacc);
            if (lane == 0) acc_add(energy + sidx * E_SLOTS + E_DIRECT, etot, ENERGY_SCALE, status);
        }
        if (STATS) {
            npair = wave_sum_u64(npair);
            hsum = wave_sum_u64(hsum);
            if (lane == 0) {
                atomicAdd(stats + 2 * sidx, npair);
                atomicAdd(stats + 2 * sidx + 1, hsum);
            }
        }
    }
}

template <class A>
__global__ void k_exceptions(int count, const int* __restrict__ p1, const int* __restrict__ p2,
                             const double* __restrict__ eqq, const double* __restrict__ esig,
                             const double* __restrict__ eeps, const double* __restrict__ charge,
                             const double* __restrict__ pos, const int* __restrict__ atom_sys,
                             const SysInfo* __restrict__ sys, int use_pbc, int has_pme, double alpha,
                             const double* __restrict__ c6, double lj_alpha, A* force, A* energy, unsigned* status) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    int si = -1;
    if (k < count) {
        int i = p1[k], j = p2[k];
        si = atom_sys[i];
        e = exception_term(i, j, eqq[k], esig[k], eeps[k], charge[i], charge[j], pos, sys[si].box, use_pbc, has_pme,
                           alpha, c6 ? c6[k] : 0.0, lj_alpha, force, status);
    }
    wave_energy_add(energy, si, E_EXCEPT, e, status);
}

template <class A>
A* slot_buffer(const NlistDev& nl);
template <>
long long* slot_buffer<long long>(const NlistDev& nl) {
    return nl.fslot_fx.p;
}
template <>
double* slot_buffer<double>(const NlistDev& nl) {
    return nl.fslot_d.p;
}

// Adds the slot-ordered pair forces into the atom-ordered accumulator and
// clears them.  Atomic, because the mesh may add into the same accumulator
// from another stream meanwhile.  order_a/order_b (nullable): copies of the
// slot order it reads, for the meshes' next run.
__device__ __forceinline__ void add_raw(long long* p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
__device__ __forceinline__ void add_raw(double* p, double v) { unsafeAtomicAdd(p, v); }

template <class A>
__global__ void k_slot_forces(const int* __restrict__ sorted_atom, int nslots, A* fslot, A* force, int* order_a,
                              int* order_b) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int a = sorted_atom[s];
    if (order_a) order_a[s] = a;
    if (order_b) order_b[s] = a;
    for (int c = 0; c < 3; ++c) {
        const A v = fslot[(size_t)c * nslots + s];
        if (v != (A)0) {
            if (a >= 0) add_raw(force + 3 * a + c, v);
            fslot[(size_t)c * nslots + s] = (A)0;
        }
    }
}

template <class R, class A, int ELEC, bool STATS, bool ENERGY>
void launch_one(const NlistDev& nl, const PairParams& p, const R* qk, const R* sh, const R* e2,
                const ParticleDev& part, const double* pos, const SysInfo* sys_dev, A* force, A* energy,
                unsigned* status, unsigned long long* stats, int grid, bool combine, int* const* order_out,
                hipStream_t s) {
    hipLaunchKernelGGL((k_pairs<R, A, ELEC, STATS, ENERGY>), dim3(grid), dim3(64), 0, s, pos, qk, sh, e2,
                       part.qk_d.p, part.sh_d.p, part.e2_d.p, nl.sorted_atom.p, nl.clist_ex.p, nl.block_center.p,
                       nl.block_minimg.p, nl.clist.p, nl.ccount.p, nl.tasks.p, nl.ntasks.p, nl.block_sys.p, sys_dev, p,
                       slot_buffer<A>(nl), energy, status, stats, nl.flag.p, nl.nsys + 1);
    if (combine)
        hipLaunchKernelGGL(k_slot_forces<A>, dim3((nl.nslots + 255) / 256), dim3(256), 0, s, nl.sorted_atom.p,
                           nl.nslots, slot_buffer<A>(nl), force, order_out[0], order_out[1]);
}

template <class R, class A, int ELEC>
void launch_pairs_t(bool stats_on, bool energy_on, const NlistDev& nl, const PairParams& p, const R* qk, const R* sh,
                    const R* e2, const ParticleDev& part, const double* pos, const SysInfo* sys_dev, A* force, A* energy,
                    unsigned* status, unsigned long long* stats, int grid, bool combine, int* const* order_out,
                    hipStream_t s) {
    if (stats_on)
        launch_one<R, A, ELEC, true, false>(nl, p, qk, sh, e2, part, pos, sys_dev, force, energy, status, stats, grid,
                                            combine, order_out, s);
    else if (energy_on)
        launch_one<R, A, ELEC, false, true>(nl, p, qk, sh, e2, part, pos, sys_dev, force, energy, status, stats, grid,
                                            combine, order_out, s);
    else
        launch_one<R, A, ELEC, false, false>(nl, p, qk, sh, e2, part, pos, sys_dev, force, energy, status, stats, grid,
                                             combine, order_out, s);
}

}  // namespace

namespace {
// Pair kernel variants: the electrostatics kind, plus PAIR_SWITCH when the LJ
// switching function is on (a separate kernel, so the common case carries
// none of its arithmetic).
template <class R, class A>
void dispatch_kind(int kind, bool st, bool energy_on, const NlistDev& nl, const PairParams& p, const R* qk,
                   const R* sh, const R* e2, const ParticleDev& part, const double* pos, const SysInfo* sys_dev,
                   A* force, A* energy, unsigned* status, unsigned long long* stats, int grid, bool combine,
                   int* const* order_out, hipStream_t s) {
    switch (kind) {
        case ELEC_EWALD:
            launch_pairs_t<R, A, ELEC_EWALD>(st, energy_on, nl, p, qk, sh, e2, part, pos, sys_dev, force, energy, status,
                                             stats, grid, combine, order_out, s);
            break;
        case ELEC_REACTION_FIELD:
            launch_pairs_t<R, A, ELEC_REACTION_FIELD>(st, energy_on, nl, p, qk, sh, e2, part, pos, sys_dev, force,
                                                      energy, status, stats, grid, combine, order_out, s);
            break;
        case ELEC_COULOMB:
            launch_pairs_t<R, A, ELEC_COULOMB>(st, energy_on, nl, p, qk, sh, e2, part, pos, sys_dev, force, energy,
                                               status, stats, grid, combine, order_out, s);
            break;
        case ELEC_EWALD | PAIR_SWITCH:
            launch_pairs_t<R, A, ELEC_EWALD | PAIR_SWITCH>(st, energy_on, nl, p, qk, sh, e2, part, pos, sys_dev, force,
                                                           energy, status, stats, grid, combine, order_out, s);
            break;
        case ELEC_REACTION_FIELD | PAIR_SWITCH:
            launch_pairs_t<R, A, ELEC_REACTION_FIELD | PAIR_SWITCH>(st, energy_on, nl, p, qk, sh, e2, part, pos,
                                                                    sys_dev, force, energy, status, stats, grid,
                                                                    combine, order_out, s);
            break;
        case ELEC_EWALD | PAIR_LJPME:
            launch_pairs_t<R, A, ELEC_EWALD | PAIR_LJPME>(st, energy_on, nl, p, qk, sh, e2, part, pos, sys_dev, force,
                                                          energy, status, stats, grid, combine, order_out, s);
            break;
        default:
            throw Error(OMM_ERR_INVALID, "no pair kernel for this electrostatics / switching combination");
    }
}
}  // namespace

void launch_pairs(bool mixed, bool energy_on, const NlistDev& nl, const PairParams& p, const ParticleDev& part,
                  const double* pos, const SysInfo* sys_dev, long long* force_fx, double* force_d, long long* energy_fx,
                  double* energy_d, unsigned* status, unsigned long long* stats, int grid, hipStream_t s,
                  bool combine, int* order_a, int* order_b) {
    const bool st = stats != nullptr;
    int* const order_out[2] = {order_a, order_b};
    const int kind = p.elec | (p.switching ? PAIR_SWITCH : 0) | (p.ljpme ? PAIR_LJPME : 0);
    if (mixed)
        dispatch_kind<float, long long>(kind, st, energy_on, nl, p, part.qk.p, part.sh.p, part.e2.p, part, pos,
                                        sys_dev, force_fx, energy_fx, status, stats, grid, combine, order_out, s);
    else
        dispatch_kind<double, double>(kind, st, energy_on, nl, p, part.qk_d.p, part.sh_d.p, part.e2_d.p, part, pos,
                                      sys_dev, force_d, energy_d, status, stats, grid, combine, order_out, s);
    OMM_HIP(hipGetLastError());
}

void launch_slot_forces(bool mixed, const NlistDev& nl, long long* force_fx, double* force_d, hipStream_t s) {
    if (mixed)
        hipLaunchKernelGGL(k_slot_forces<long long>, dim3((nl.nslots + 255) / 256), dim3(256), 0, s, nl.sorted_atom.p,
                           nl.nslots, nl.fslot_fx.p, force_fx, nullptr, nullptr);
    else
        hipLaunchKernelGGL(k_slot_forces<double>, dim3((nl.nslots + 255) / 256), dim3(256), 0, s, nl.sorted_atom.p,
                           nl.nslots, nl.fslot_d.p, force_d, nullptr, nullptr);
    OMM_HIP(hipGetLastError());
}

void launch_exceptions(bool mixed, const ExceptionDev& ex, const ParticleDev& part, const double* pos,
                       const SysInfo* sys_dev, int use_pbc, int has_pme, double alpha, double lj_alpha,
                       long long* force_fx, double* force_d, long long* energy_fx, double* energy_d, unsigned* status,
                       hipStream_t s) {
    if (ex.count == 0) return;
    int blocks = (ex.count + 255) / 256;
    const double* c6 = lj_alpha > 0.0 ? ex.c6.p : nullptr;
    if (mixed)
        hipLaunchKernelGGL(k_exceptions<long long>, dim3(blocks), dim3(256), 0, s, ex.count, ex.p1.p, ex.p2.p, ex.qq.p,
                           ex.sig.p, ex.eps.p, part.charge.p, pos, part.atom_sys.p, sys_dev, use_pbc, has_pme, alpha,
                           c6, lj_alpha, force_fx, energy_fx, status);
    else
        hipLaunchKernelGGL(k_exceptions<double>, dim3(blocks), dim3(256), 0, s, ex.count, ex.p1.p, ex.p2.p, ex.qq.p,
                           ex.sig.p, ex.eps.p, part.charge.p, pos, part.atom_sys.p, sys_dev, use_pbc, has_pme, alpha,
                           c6, lj_alpha, force_d, energy_d, status);
    OMM_HIP(hipGetLastError());
}

}  // namespace ommhip
