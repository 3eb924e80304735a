// Smooth PME on the device: B-spline spread into a fixed-point (mixed) or
// FP64 (plain) grid, a real-to-complex 3D FFT as three passes of batched 1D
// mixed-radix Stockham transforms held in LDS, the reciprocal convolution and
// its energy fused into the x pass (forward x, multiply, inverse x in one
// kernel), the inverse passes and the force gather.
//
// The reference transforms the full complex grid (pme.rs:271-303) and sums the
// energy over all of it (pme.rs:339-372); for a real charge grid S(-k) =
// conj(S(k)) and the influence function is even, so the half spectrum
// kz <= nz/2 with weights 1 (kz = 0, nz/2) and 2 (otherwise) carries the same
// sum.  Any grid size is supported: radices are the prime factors of n (8 and
// 4 preferred), each stage a direct DFT of its radix.
#include <climits>
#include <cmath>
#include <complex>
#include <type_traits>

#include "fmath.h"
#include "pme.h"
#include "spline.h"

namespace ommhip {

namespace {

constexpr int FFT_THREADS = 256;

template <class R>
struct Cx {
    R x, y;
};

// Spread without atomics on the grid.  The grid of every system is cut into tiles
// of ptile = 8 or 6 points per axis (the last tile of an axis takes the
// remainder, so tile widths lie in [4, 11]; smaller tiles give small grids
// more workgroups).  Each step, k_pme_bin computes every charged atom's
// stencil (base point and weights, stored per atom) and appends the atom to
// the bin of the tile holding its base point.  An atom's 5-point stencil then
// only reaches its own tile and the next one along each axis, so the points
// of tile T receive contributions from the bins of T and of the tiles before
// it: at most 8 bins.  k_spread_tiles runs one workgroup per tile, sums those
// contributions in LDS (fixed point: the order of the adds does not matter)
// and writes every point of the tile with plain stores.  Each point has one
// owner, and the whole grid is overwritten.
constexpr int TILE_MAX = 11;
constexpr int SPREAD_THREADS = 256;
constexpr int SPREAD_MIN_TILES = 1024;  // smaller tile edges until a batch has this many tiles

__device__ __forceinline__ int tile_of(int i, int tiles, int edge) { return min(i / edge, tiles - 1); }

// Tiles along an axis of n points for edge e: widths e ... (last) in [4, 11].
inline int tiles_for(int n, int e) { return e == 8 ? std::max(1, (n + 4) / 8) : std::max(1, n / e); }

// Atoms are visited in the neighbour list's sorted order, so the lanes of a
// wave mostly fall into one or two tiles, or (order == nullptr) in atom
// order, which lets the spread start before a list build has placed the
// atoms; each distinct tile of a wave takes one atomic for all of its lanes
// (contended device-scope atomics serialise).  The bins' order never changes
// a result: the spread's fixed-point sums are order-independent.
template <class R>
__global__ void k_pme_bin(int nslots, const int* __restrict__ order, const double* __restrict__ pos,
                          const double* __restrict__ charge, const int* __restrict__ atom_sys,
                          const SysInfo* __restrict__ sys, int* st_idx, R* st_w, int* bin_count, int* bins,
                          int bin_cap, unsigned* status) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int a = slot < nslots ? (order ? order[slot] : slot) : -1;
    const double q = a >= 0 ? charge[a] : 0.0;
    int tile = -1;
    if (q != 0.0) {  // uncharged atoms do not spread (pme.rs:315-330)
        const SysInfo& S = sys[atom_sys[a]];
        int idx[3];
        R w[3][5], dw[3][5];
        stencil<R>(pos + 3 * a, S.box, S.grid, idx, w, dw);
        for (int d = 0; d < 3; ++d) st_idx[4 * a + d] = idx[d];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 5; ++k) st_w[16 * a + 5 * d + k] = w[d][k];
        tile = S.tile_off +
               (tile_of(idx[0], S.tiles[0], S.ptile) * S.tiles[1] + tile_of(idx[1], S.tiles[1], S.ptile)) *
                   S.tiles[2] +
               tile_of(idx[2], S.tiles[2], S.ptile);
    }
    unsigned long long pending = __ballot(tile >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int t = __shfl(tile, leader, 64);
        const unsigned long long group = __ballot(tile == t);
        int base = 0;
        if (lane == leader) base = atomicAdd(&bin_count[t], __popcll(group));
        base = __shfl(base, leader, 64);
        if (tile == t) {
            const int p = base + __popcll(group & ((1ull << lane) - 1ull));
            if (p < bin_cap) bins[(size_t)t * bin_cap + p] = a;
            else atomicOr(status, ST_BIN_OVERFLOW);
        }
        pending &= ~group;
    }
}

__device__ __forceinline__ void lds_add(unsigned long long* p, long long v) {
    if (v) atomicAdd(p, (unsigned long long)v);
}
__device__ __forceinline__ void lds_add(double* p, double v) {
    if (v != 0.0) atomicAdd(p, v);
}

// FP32 spread into the fixed-point grid: the contributions of one staged
// round of SPREAD_THREADS atoms are summed as int32 at 2^-24 and added to the
// int64 tile sums at the end of the round.  A contribution q w is at most
// |q| 0.2153 (the largest 3-D weight of the order-5 spline is (115/192)^3),
// so a round's sum stays below 2^31 / 2^24 = 128 whenever |q| <= 2 for every
// atom (SPREAD32_MAX_CHARGE); otherwise the int64 path is used throughout.
constexpr double SPREAD32_SCALE = 16777216.0;  // 2^24
constexpr double SPREAD32_MAX_CHARGE = 2.0;

template <class R, class G>
__global__ __launch_bounds__(SPREAD_THREADS) void k_spread_tiles(const int* __restrict__ tile_off, int nsys,
                                                                 const SysInfo* __restrict__ sys,
                                                                 const int* __restrict__ st_idx,
                                                                 const R* __restrict__ st_w,
                                                                 const double* __restrict__ charge,
                                                                 const int* __restrict__ bin_count,
                                                                 const int* __restrict__ bins, int bin_cap, G* grid,
                                                                 int small_charges, unsigned* status) {
    constexpr bool FIXED = sizeof(G) == sizeof(long long) && G(0.5) == G(0);  // int64 grid
    using Acc = typename std::conditional<FIXED, unsigned long long, double>::type;
    constexpr bool I32 = FIXED && sizeof(R) == 4;
    const bool use32 = I32 && small_charges;
    __shared__ Acc s_acc[TILE_MAX * TILE_MAX * TILE_MAX];
    __shared__ int s_acc32[I32 ? TILE_MAX * TILE_MAX * TILE_MAX : 1];
    __shared__ int s_src[8], s_pre[9];
    __shared__ int s_idx[3][SPREAD_THREADS];
    __shared__ R s_w[15][SPREAD_THREADS];
    __shared__ double s_q[SPREAD_THREADS];
    const int si = find_sys(tile_off, nsys, blockIdx.x);
    const SysInfo& S = sys[si];
    const int T0 = S.tiles[0], T1 = S.tiles[1], T2 = S.tiles[2];
    const int local = blockIdx.x - S.tile_off;
    const int tz = local % T2, txy = local / T2;
    const int ty = txy % T1, tx = txy / T1;
    const int tcoord[3] = {tx, ty, tz}, tiles[3] = {T0, T1, T2};
    int lo[3], wid[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = tcoord[d] * S.ptile;
        wid[d] = (tcoord[d] == tiles[d] - 1 ? S.grid[d] : lo[d] + S.ptile) - lo[d];
    }
    const int npts = wid[0] * wid[1] * wid[2];
    for (int i = threadIdx.x; i < npts; i += SPREAD_THREADS) {
        s_acc[i] = (Acc)0;
        if (I32) s_acc32[i] = 0;
    }
    if (threadIdx.x == 0) {
        // Source bins: this tile and the previous one along each axis
        // (periodic), without repeats when an axis has one or two tiles.
        int nsrc = 0;
        for (int dx = -1; dx <= 0; ++dx)
            for (int dy = -1; dy <= 0; ++dy)
                for (int dz = -1; dz <= 0; ++dz) {
                    if ((dx && T0 == 1) || (dy && T1 == 1) || (dz && T2 == 1)) continue;
                    const int sx = (tx + dx + T0) % T0, sy = (ty + dy + T1) % T1, sz = (tz + dz + T2) % T2;
                    s_src[nsrc++] = S.tile_off + (sx * T1 + sy) * T2 + sz;
                }
        s_pre[0] = 0;
        for (int b = 0; b < nsrc; ++b) s_pre[b + 1] = s_pre[b] + min(bin_count[s_src[b]], bin_cap);
        for (int b = nsrc; b < 8; ++b) s_pre[b + 1] = s_pre[b];
    }
    __syncthreads();
    // Source atoms are staged in LDS, SPREAD_THREADS at a time, with one
    // independent load chain per thread; the work items are then (staged
    // atom, x plane of its stencil), five per atom, so atoms that reach only
    // a corner of the tile do not hold up the others.
    const int natoms_src = s_pre[8];
    for (int base = 0; base < natoms_src; base += SPREAD_THREADS) {
        const int k = base + threadIdx.x;
        if (k < natoms_src) {
            int b = 0;
            while (s_pre[b + 1] <= k) ++b;
            const int a = bins[(size_t)s_src[b] * bin_cap + (k - s_pre[b])];
            for (int d = 0; d < 3; ++d) s_idx[d][threadIdx.x] = st_idx[4 * a + d];
            for (int j = 0; j < 15; ++j) s_w[j][threadIdx.x] = st_w[16 * (size_t)a + j];
            s_q[threadIdx.x] = charge[a];
        }
        __syncthreads();
        const int cnt = min(SPREAD_THREADS, natoms_src - base);
        for (int item = threadIdx.x; item < 5 * cnt; item += SPREAD_THREADS) {
            const int m = item / 5, ix = item - 5 * m;
            int px = s_idx[0][m] + ix;
            if (px >= S.grid[0]) px -= S.grid[0];
            const int lx = px - lo[0];
            if (lx < 0 || lx >= wid[0]) continue;
            const double q = s_q[m];
            int loc[2][5];
            for (int d = 0; d < 2; ++d) {
                const int n = S.grid[d + 1];
                int p = s_idx[d + 1][m];
                for (int j = 0; j < 5; ++j) {
                    const int l = p - lo[d + 1];
                    loc[d][j] = (l >= 0 && l < wid[d + 1]) ? l : -1;
                    if (++p == n) p = 0;
                }
            }
            const R wx = s_w[ix][m];
            if (I32 && use32) {
                const float qx = (float)(q * SPREAD32_SCALE) * (float)wx;
                for (int iy = 0; iy < 5; ++iy) {
                    if (loc[0][iy] < 0) continue;
                    const float qxy = qx * (float)s_w[5 + iy][m];
                    int* row = s_acc32 + (lx * wid[1] + loc[0][iy]) * wid[2];
                    for (int iz = 0; iz < 5; ++iz) {
                        if (loc[1][iz] < 0) continue;
                        const int v = __float2int_rn(qxy * (float)s_w[10 + iz][m]);
                        if (v) atomicAdd(row + loc[1][iz], v);
                    }
                }
                continue;
            }
            for (int iy = 0; iy < 5; ++iy) {
                if (loc[0][iy] < 0) continue;
                const R wxy = wx * s_w[5 + iy][m];
                Acc* row = s_acc + (lx * wid[1] + loc[0][iy]) * wid[2];
                for (int iz = 0; iz < 5; ++iz) {
                    if (loc[1][iz] < 0) continue;
                    const double c = q * (double)(wxy * s_w[10 + iz][m]);
                    if (FIXED) lds_add(row + loc[1][iz], (long long)to_fixed(c, GRID_SCALE, status));
                    else lds_add(row + loc[1][iz], c);
                }
            }
        }
        __syncthreads();
        if (I32 && use32) {
            // Fold the round's int32 sums into the int64 tile sums (2^-24 -> 2^-40).
            for (int i = threadIdx.x; i < npts; i += SPREAD_THREADS) {
                s_acc[i] += (Acc)((long long)s_acc32[i] * 65536ll);
                s_acc32[i] = 0;
            }
            __syncthreads();
        }
    }
    G* g = grid + S.qgrid_off;
    const int ny = S.grid[1], nz = S.grid[2];
    for (int i = threadIdx.x; i < npts; i += SPREAD_THREADS) {
        const int lz = i % wid[2], r = i / wid[2];
        const int ly = r % wid[1], lx = r / wid[1];
        g[((size_t)(lo[0] + lx) * ny + (lo[1] + ly)) * nz + lo[2] + lz] = (G)s_acc[i];
    }
}

// Atom-owned spread (the autotuner's alternative to the tile-owned one): one
// workgroup per chunk of PME_CHUNK sorted slots, or of PME_CHUNK atoms in
// index order while no sorted order is available.  The chunk's stencils are
// computed once.  When the chunk lies in one system and its stencil points
// fit a box of at most SPREAD_BOX_BYTES of LDS, the contributions are summed
// in that box and the box is added to the grid with one device atomic per
// non-zero point; otherwise every contribution is added to the grid
// directly.  Each contribution is the tile-owned spread's fixed-point value,
// and integer sums do not depend on their order, so both spreads write the
// same bits.  The grid is zero at the launch (the forward transform after
// the previous spread cleared every point it read).
constexpr int SPREAD_BOX_BYTES = 32768;
constexpr int ATOM_SPREAD_THREADS = 256;
static_assert(PME_CHUNK == 64, "the atom-owned spread computes a chunk's stencils in one wave");

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wrap_index(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}
__device__ __forceinline__ void grid_add(long long* p, unsigned long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), v);
}
__device__ __forceinline__ void grid_add(double* p, double v) { atomicAdd(p, v); }

template <class R, class G>
__global__ __launch_bounds__(ATOM_SPREAD_THREADS) void k_spread_atoms(const int* __restrict__ chunk_off, int nsys,
                                                                      const int* __restrict__ order, int natoms,
                                                                      const double* __restrict__ pos,
                                                                      const double* __restrict__ charge,
                                                                      const int* __restrict__ atom_sys,
                                                                      const SysInfo* __restrict__ sys, G* grid,
                                                                      int small_charges, unsigned* status) {
    constexpr bool FIXED = sizeof(G) == sizeof(long long) && G(0.5) == G(0);  // int64 grid
    using Acc = typename std::conditional<FIXED, unsigned long long, double>::type;
    constexpr bool I32 = FIXED && sizeof(R) == 4;
    const bool use32 = I32 && small_charges;
    // The box: int32 sums at 2^-24 (at most PME_CHUNK contributions of
    // |q| w <= 2 x 0.2153 per point stay below 2^31) or Acc sums.
    __shared__ __align__(8) unsigned char s_raw[SPREAD_BOX_BYTES];
    int* box32 = reinterpret_cast<int*>(s_raw);
    Acc* box = reinterpret_cast<Acc*>(s_raw);
    __shared__ int s_base[3][PME_CHUNK];  // box-local (fit) or grid base point
    __shared__ R s_w[15][PME_CHUNK];
    __shared__ double s_q[PME_CHUNK];
    __shared__ int s_sys[PME_CHUNK];
    __shared__ int s_org[3], s_ext[3], s_fit, s_lsys;
    const int t = threadIdx.x;
    if (t < PME_CHUNK) {
        int a = -1;
        if (chunk_off) {
            const SysInfo& S = sys[find_sys(chunk_off, nsys, blockIdx.x)];
            const int slot = S.slot_off + (blockIdx.x - S.chunk_off) * PME_CHUNK + t;
            a = slot < S.slot_off + S.nblocks * IBLOCK ? order[slot] : -1;
        } else {
            const int i = blockIdx.x * PME_CHUNK + t;
            a = i < natoms ? i : -1;
        }
        const double q = a >= 0 ? charge[a] : 0.0;
        const bool on = q != 0.0;  // uncharged atoms do not spread (pme.rs:315-330)
        int si = 0, idx[3] = {0, 0, 0};
        if (on) {
            si = atom_sys[a];
            const SysInfo& S = sys[si];
            R w[3][5], dw[3][5];
            stencil<R>(pos + 3 * a, S.box, S.grid, idx, w, dw);
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 5; ++k) s_w[5 * d + k][t] = w[d][k];
        }
        s_q[t] = q;
        s_sys[t] = si;
        const unsigned long long live = __ballot(on);
        int fit = -1;  // nothing to spread
        if (live) {
            const int lead = __ffsll((long long)live) - 1;
            const int lsys = __shfl(si, lead, 64);
            const SysInfo& L = sys[lsys];
            int rel[3], lo[3], ext[3];
            for (int d = 0; d < 3; ++d) {
                // Offsets from the lead atom's base point, taken in (-n/2, n/2].
                const int n = L.grid[d];
                int r = idx[d] - __shfl(idx[d], lead, 64);
                if (r > n / 2) r -= n;
                else if (r <= -(n - n / 2)) r += n;
                rel[d] = r;
                lo[d] = wave_min(on ? r : INT_MAX);
                ext[d] = wave_max(on ? r : INT_MIN) - lo[d] + 5;
            }
            const long long vol = (long long)ext[0] * ext[1] * ext[2];
            const int cap = SPREAD_BOX_BYTES / (use32 ? (int)sizeof(int) : (int)sizeof(Acc));
            fit = (__ballot(on && si != lsys) == 0 && vol <= cap) ? 1 : 0;
            for (int d = 0; d < 3; ++d) s_base[d][t] = fit ? rel[d] - lo[d] : idx[d];
            if (t == 0) {
                for (int d = 0; d < 3; ++d) {
                    s_org[d] = wrap_index(__shfl(idx[d], lead, 64) + lo[d], L.grid[d]);
                    s_ext[d] = ext[d];
                }
                s_lsys = lsys;
            }
        }
        if (t == 0) s_fit = fit;
    }
    __syncthreads();
    const int fit = s_fit;
    if (fit < 0) return;
    if (fit) {
        const int e1 = s_ext[1], e2 = s_ext[2];
        const int vol = s_ext[0] * e1 * e2;
        for (int i = t; i < vol; i += ATOM_SPREAD_THREADS) {
            if (I32 && use32) box32[i] = 0;
            else box[i] = (Acc)0;
        }
        __syncthreads();
        for (int item = t; item < PME_CHUNK * 25; item += ATOM_SPREAD_THREADS) {
            const int m = item / 25, r = item - 25 * m;
            const int ix = r / 5, iy = r - 5 * ix;
            const double q = s_q[m];
            if (q == 0.0) continue;
            const int row = ((s_base[0][m] + ix) * e1 + s_base[1][m] + iy) * e2 + s_base[2][m];
            const R wx = s_w[ix][m];
            if (I32 && use32) {
                const float qx = (float)(q * SPREAD32_SCALE) * (float)wx;
                const float qxy = qx * (float)s_w[5 + iy][m];
                for (int iz = 0; iz < 5; ++iz) {
                    const int v = __float2int_rn(qxy * (float)s_w[10 + iz][m]);
                    if (v) atomicAdd(box32 + row + iz, v);
                }
            } else {
                const R wxy = wx * s_w[5 + iy][m];
                for (int iz = 0; iz < 5; ++iz) {
                    const double c = q * (double)(wxy * s_w[10 + iz][m]);
                    if (FIXED) lds_add(reinterpret_cast<unsigned long long*>(box) + row + iz,
                                       (long long)to_fixed(c, GRID_SCALE, status));
                    else lds_add(reinterpret_cast<double*>(box) + row + iz, c);
                }
            }
        }
        __syncthreads();
        const SysInfo& L = sys[s_lsys];
        const int nx = L.grid[0], ny = L.grid[1], nz = L.grid[2];
        G* g = grid + L.qgrid_off;
        for (int i = t; i < vol; i += ATOM_SPREAD_THREADS) {
            Acc v;
            if (I32 && use32) {
                const int v32 = box32[i];
                if (!v32) continue;
                v = (Acc)((long long)v32 * 65536ll);
            } else {
                v = box[i];
                if (v == (Acc)0) continue;
            }
            const int lz = i % e2, rr = i / e2;
            const int ly = rr % e1, lx = rr / e1;
            const int gx = wrap_index(s_org[0] + lx, nx), gy = wrap_index(s_org[1] + ly, ny),
                      gz = wrap_index(s_org[2] + lz, nz);
            grid_add(g + ((size_t)gx * ny + gy) * nz + gz, v);
        }
        return;
    }
    // Direct adds (a chunk spread over more than a box, or over systems).
    for (int item = t; item < PME_CHUNK * 25; item += ATOM_SPREAD_THREADS) {
        const int m = item / 25, r = item - 25 * m;
        const int ix = r / 5, iy = r - 5 * ix;
        const double q = s_q[m];
        if (q == 0.0) continue;
        const SysInfo& S = sys[s_sys[m]];
        const int nx = S.grid[0], ny = S.grid[1], nz = S.grid[2];
        int gx = s_base[0][m] + ix, gy = s_base[1][m] + iy;
        if (gx >= nx) gx -= nx;
        if (gy >= ny) gy -= ny;
        G* grow = grid + S.qgrid_off + ((size_t)gx * ny + gy) * nz;
        const int gz0 = s_base[2][m];
        const R wx = s_w[ix][m];
        if (I32 && use32) {
            const float qx = (float)(q * SPREAD32_SCALE) * (float)wx;
            const float qxy = qx * (float)s_w[5 + iy][m];
            for (int iz = 0; iz < 5; ++iz) {
                const int v = __float2int_rn(qxy * (float)s_w[10 + iz][m]);
                int gz = gz0 + iz;
                if (gz >= nz) gz -= nz;
                if (v) grid_add(grow + gz, (unsigned long long)((long long)v * 65536ll));
            }
        } else {
            const R wxy = wx * s_w[5 + iy][m];
            for (int iz = 0; iz < 5; ++iz) {
                const double c = q * (double)(wxy * s_w[10 + iz][m]);
                int gz = gz0 + iz;
                if (gz >= nz) gz -= nz;
                if (FIXED) {
                    const long long f = to_fixed(c, GRID_SCALE, status);
                    if (f) grid_add(grow + gz, (unsigned long long)f);
                } else if (c != 0.0) {
                    grid_add(grow + gz, c);
                }
            }
        }
    }
}

// Gather with five threads per atom (one per x plane of the stencil), atoms
// in the spatially sorted order; the five partial gradients are combined in
// LDS.
constexpr int GATHER_ATOMS = PME_CHUNK;

template <class R, class A>
__global__ __launch_bounds__(GATHER_ATOMS * 5) void k_gather5(const int* __restrict__ chunk_off, int nsys,
                                                              const int* __restrict__ order,
                                                              const double* __restrict__ pos,
                                                              const double* __restrict__ charge,
                                                              const SysInfo* __restrict__ sys,
                                                              const R* __restrict__ phi, A* force, unsigned* status) {
    __shared__ R s_g[GATHER_ATOMS][5][3];
    const int si = find_sys(chunk_off, nsys, blockIdx.x);
    const SysInfo& S = sys[si];
    const int nx = S.grid[0], ny = S.grid[1], nz = S.grid[2];
    phi += S.qgrid_off;
    const int t = threadIdx.x;
    const int k = t / 5, ix = t - k * 5;
    const int slot = S.slot_off + (blockIdx.x - S.chunk_off) * GATHER_ATOMS + k;
    const int a = slot < S.slot_off + S.nblocks * IBLOCK ? order[slot] : -1;
    const double q = a >= 0 ? charge[a] : 0.0;
    const Box& b = S.box;
    R g0 = 0, g1 = 0, g2 = 0;
    if (q != 0.0) {
        const int dims[3] = {nx, ny, nz};
        int idx[3];
        R w[3][5], dw[3][5];
        stencil<R>(pos + 3 * a, b, dims, idx, w, dw);
        int gx = idx[0] + ix;
        if (gx >= nx) gx -= nx;
        int zz[5];
        for (int iz = 0; iz < 5; ++iz) {
            zz[iz] = idx[2] + iz;
            if (zz[iz] >= nz) zz[iz] -= nz;
        }
        for (int iy = 0; iy < 5; ++iy) {
            int gy = idx[1] + iy;
            if (gy >= ny) gy -= ny;
            const R wxy = w[0][ix] * w[1][iy];
            const R d0 = dw[0][ix] * w[1][iy];
            const R d1 = w[0][ix] * dw[1][iy];
            const R* row = phi + ((size_t)gx * ny + gy) * nz;
            for (int iz = 0; iz < 5; ++iz) {
                const R p = row[zz[iz]];
                g0 += p * d0 * w[2][iz];
                g1 += p * d1 * w[2][iz];
                g2 += p * wxy * dw[2][iz];
            }
        }
    }
    s_g[k][ix][0] = g0;
    s_g[k][ix][1] = g1;
    s_g[k][ix][2] = g2;
    __syncthreads();
    if (ix == 0 && q != 0.0) {
        R G0 = 0, G1 = 0, G2 = 0;
        for (int m = 0; m < 5; ++m) {
            G0 += s_g[k][m][0];
            G1 += s_g[k][m][1];
            G2 += s_g[k][m][2];
        }
        // pme.rs:417-422
        for (int c = 0; c < 3; ++c) {
            double s = (double)G0 * nx * b.recip[0 * 3 + c] + (double)G1 * ny * b.recip[1 * 3 + c] +
                       (double)G2 * nz * b.recip[2 * 3 + c];
            acc_add(force + 3 * a + c, -q * s, FORCE_SCALE, status);
        }
    }
}

enum : int { MODE_Z_R2C = 0, MODE_Y_FWD = 1, MODE_X_CONV = 2, MODE_Y_INV = 3, MODE_Z_C2R = 4 };


__device__ __forceinline__ double grid_real(long long v) { return (double)v * (1.0 / GRID_SCALE); }
__device__ __forceinline__ double grid_real(double v) { return v; }

// x / d for 0 <= x < 2^20 and 1 <= d: a single-precision estimate corrected
// by one step either way.
__device__ __forceinline__ int div_small(int x, int d, float inv_d) {
    int q = (int)((float)x * inv_d);
    if (q * d > x) --q;
    else if ((q + 1) * d <= x) ++q;
    return q;
}

// Batched Stockham FFT of `lines` lines of length n held line-major in a/b,
// T threads per line.  A stage of radix R is a direct DFT per output:
// out[(j - k) R + k + q Ns] = sum_r in[j + r m] W_n^{r (k n/(Ns R) + q m)},
// m = n/R, output index o = q m + j, k = j mod Ns.  The exponent base
// k n/(Ns R) + q m stays below n, so the root index wraps with one conditional
// subtraction.  Each output is computed by one thread in a fixed order, so
// the result does not depend on T or on how lines are grouped.
__device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class R>
__device__ Cx<R>* fft_lds(Cx<R>* a, Cx<R>* b, int n, const int* radices, const Cx<R>* roots, bool inverse, int T) {
    const int li = threadIdx.x / T, t = threadIdx.x - li * T;
    const R sgn = inverse ? (R)-1 : (R)1;
    int Ns = 1;
    for (int s = 0; s < MAX_STAGES && radices[s]; ++s) {
        const int Rr = radices[s];
        const int m = n / Rr;
        const int step = n / (Ns * Rr);
        const Cx<R>* in = a + li * n;
        Cx<R>* out = b + li * n;
        const float inv_ns = 1.0f / (float)Ns;
        const float inv_m = 1.0f / (float)m;
        for (int o = t; o < n; o += T) {
            const int q = div_small(o, m, inv_m);
            const int j = o - q * m;
            const int k = j - div_small(j, Ns, inv_ns) * Ns;
            const int base = k * step + q * m;
            R ax = 0, ay = 0;
            int idx = 0;
            for (int r = 0; r < Rr; ++r) {
                const Cx<R> w = roots[idx];
                const R wy = sgn * w.y;
                const Cx<R> v = in[j + r * m];
                // Explicit fused multiply-adds: every kernel that inlines
                // this routine rounds the same way.
                ax = fused(v.x, w.x, fused(-v.y, wy, ax));
                ay = fused(v.x, wy, fused(v.y, w.x, ay));
                idx += base;
                if (idx >= n) idx -= n;
            }
            out[(j - k) * Rr + k + q * Ns] = Cx<R>{ax, ay};
        }
        __syncthreads();
        Cx<R>* tmp = a;
        a = b;
        b = tmp;
        Ns *= Rr;
    }
    return a;
}

// The z passes transform two real lines y = 2p, 2p + 1 of an x plane at once
// as the complex line a + i b (the last line alone when ny is odd).  From
// Z = FFT(a + i b) at k and at (n - k) mod n, the half spectra are
//   A = (Z_k + conj(Z_{n-k})) / 2,  B = (Z_k - conj(Z_{n-k})) / (2i);
// the inverse transforms A_full + i B_full (each rebuilt from its half by
// Hermitian symmetry) and reads a and b as the real and imaginary parts.
// Every kernel pairs the lines the same way and calls the same routines, so
// the plane and line passes keep giving the same bits.
template <class R>
__device__ __forceinline__ void unpack_pair(Cx<R> U, Cx<R> W, Cx<R>& A, Cx<R>& B) {
    A = Cx<R>{(R)0.5 * (U.x + W.x), (R)0.5 * (U.y - W.y)};
    B = Cx<R>{(R)0.5 * (U.y + W.y), (R)0.5 * (W.x - U.x)};
}
template <class R>
__device__ __forceinline__ Cx<R> pack_pair(Cx<R> A, Cx<R> B) {
    return Cx<R>{A.x - B.y, A.y + B.x};
}
// Element el of the full spectrum of a real line from its half h[0 .. nz/2].
template <class R>
__device__ __forceinline__ Cx<R> hermitian_at(const Cx<R>* h, int el, int nz, int nzh) {
    if (el < nzh) return h[el];
    const Cx<R> v = h[nz - el];
    return Cx<R>{v.x, -v.y};
}

template <class R, class G, class A, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft(const int* __restrict__ blk_off, int nsys,
                                                     const SysInfo* __restrict__ sys, G* __restrict__ qgrid,
                                                     Cx<R>* cgrid, R* phi, const Cx<R>* __restrict__ roots_all,
                                                     const R* __restrict__ moduli_all, A* energy, unsigned* status,
                                                     int* clear, int nclear) {
    extern __shared__ unsigned char smem_raw[];
    // The first pass also clears the spread's bin counts (every spread
    // workgroup is done by now), so no fill kernel opens the next spread.
    if (MODE == MODE_Z_R2C && clear)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nclear; i += gridDim.x * blockDim.x) clear[i] = 0;
    const bool zmode = MODE == MODE_Z_R2C || MODE == MODE_Z_C2R;
    const bool ymode = MODE == MODE_Y_FWD || MODE == MODE_Y_INV;
    const int axis = zmode ? 2 : (ymode ? 1 : 0);
    const int si = find_sys(blk_off, nsys, blockIdx.x);
    const SysInfo& S = sys[si];
    const int nx = S.grid[0], ny = S.grid[1], nz = S.grid[2], nzh = S.nzh;
    const int n = S.grid[axis];
    const int LINES = S.fft_lines[sizeof(R) == 8 ? 1 : 0][axis];
    const int T = FFT_THREADS / LINES;
    const int nyp = (ny + 1) / 2;  // z passes: line pairs per x plane
    const int nlines = zmode ? nx * nyp : (ymode ? nx * nzh : ny * nzh);
    const int line0 = (blockIdx.x - blk_off[si]) * LINES;
    qgrid += S.qgrid_off;
    phi += S.qgrid_off;
    cgrid += S.cgrid_off;

    Cx<R>* bufA = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* bufB = bufA + LINES * n;
    Cx<R>* roots = bufB + LINES * n;
    __shared__ int s_radix[MAX_STAGES];
    if (threadIdx.x < MAX_STAGES) s_radix[threadIdx.x] = S.radix[axis][threadIdx.x];
    for (int i = threadIdx.x; i < n; i += FFT_THREADS) roots[i] = roots_all[S.roots_off[axis] + i];

    // Load.
    for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
        int li, el;
        if (zmode) {
            li = e / n;
            el = e - li * n;
        } else {
            li = e % LINES;
            el = e / LINES;
        }
        const int L = line0 + li;
        Cx<R> v = {0, 0};
        if (L < nlines) {
            if (zmode) {
                const int x = L / nyp, y = 2 * (L - x * nyp);
                const size_t row = (size_t)x * ny + y;
                if (MODE == MODE_Z_R2C) {
                    // The grid is zero between runs: each point is cleared as
                    // it is read.
                    G& g0 = qgrid[row * nz + el];
                    v.x = (R)grid_real(g0);
                    g0 = (G)0;
                    if (y + 1 < ny) {
                        G& g1 = qgrid[(row + 1) * nz + el];
                        v.y = (R)grid_real(g1);
                        g1 = (G)0;
                    }
                } else {
                    const Cx<R> a = hermitian_at(cgrid + row * nzh, el, nz, nzh);
                    const Cx<R> b = y + 1 < ny ? hermitian_at(cgrid + (row + 1) * nzh, el, nz, nzh) : Cx<R>{0, 0};
                    v = pack_pair(a, b);
                }
            } else if (ymode) {
                const int x = L / nzh, kz = L - x * nzh;
                v = cgrid[((size_t)x * ny + el) * nzh + kz];
            } else {
                const int y = L / nzh, kz = L - y * nzh;
                v = cgrid[((size_t)el * ny + y) * nzh + kz];
            }
        }
        bufA[li * n + el] = v;
    }
    __syncthreads();
    const int* rad = s_radix;
    const bool inverse_first = MODE == MODE_Y_INV || MODE == MODE_Z_C2R;
    Cx<R>* res = fft_lds<R>(bufA, bufB, n, rad, roots, inverse_first, T);

    if (MODE == MODE_X_CONV) {
        // Reciprocal convolution and energy (pme.rs:332-372) on the half spectrum.
        const Box& b = S.box;
        const R* mod_x = moduli_all + S.moduli_off[0];
        const R* mod_y = moduli_all + S.moduli_off[1];
        const R* mod_z = moduli_all + S.moduli_off[2];
        const double pi2_a2 = PI * PI / (S.alpha * S.alpha);
        const double pref = ONE_4PI_EPS0 / (PI * b.volume);
        // Dispersion mesh (pme.rs:358-362): I(k) = -pi^1.5 / (3 V B(k))
        // (2 pi^3 sqrt(pi) erfc(b) k^3 + e^{-b^2} (alpha^3 - 2 alpha pi^2 k^2)),
        // b = pi k / alpha, with the k = 0 term kept.
        const bool disp = S.dispersion != 0;
        const double dpref = -PI * sqrt(PI) / (3.0 * b.volume);
        const double dc1 = 2.0 * PI * PI * PI * sqrt(PI), da3 = S.alpha * S.alpha * S.alpha;
        const double dc2 = 2.0 * S.alpha * PI * PI, pi_a = PI / S.alpha;
        // The FP32 mesh serves force-only evaluations: the influence function
        // in FP32 (exp2 of the hardware), no energy.  The FP64 meshes evaluate
        // it in FP64 and sum the energy.
        //
        // The half spectrum stands for each point k and its mirror -k.  The
        // reference sums the influence over the full grid and keeps the real
        // part of the inverse transform, which is the half spectrum weighted
        // by the mean influence of k and its mirror.  The mirror's frequency
        // is -m on every axis except at a Nyquist index (2 k = n), where it is
        // m itself; in a triclinic box |k|^2 then differs between the two, so
        // those points take the mean of both influences.
        constexpr bool F32 = sizeof(R) == 4;
        float rf[9];
        for (int k = 0; k < 9; ++k) rf[k] = (float)b.recip[k];
        const bool tric = b.triclinic != 0;
        const float pi2_a2_log2e = (float)(pi2_a2 * 1.4426950408889634);
        const float pref_f = (float)pref;
        double eacc = 0.0;
        for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
            const int li = e % LINES, kx = e / LINES;
            const int L = line0 + li;
            if (L >= nlines) continue;
            const int y = L / nzh, kz = L - y * nzh;
            Cx<R>& S = res[li * n + kx];
            if (kx == 0 && y == 0 && kz == 0 && !disp) {
                S = Cx<R>{0, 0};
                continue;
            }
            // Mirror signs: the mirror's frequency is -m except at a Nyquist index.
            const bool nyq = tric && (2 * kx == nx || 2 * y == ny || 2 * kz == nz);
            const int sx = 2 * kx == nx ? 1 : -1, sy = 2 * y == ny ? 1 : -1, sz = 2 * kz == nz ? 1 : -1;
            if (F32) {
                const float mx = (float)(kx < (nx + 1) / 2 ? kx : kx - nx);
                const float my = (float)(y < (ny + 1) / 2 ? y : y - ny);
                const float mz = (float)(kz < (nz + 1) / 2 ? kz : kz - nz);
                const float mod = (float)mod_x[kx] * (float)mod_y[y] * (float)mod_z[kz];
                auto influence = [&](float ax, float ay, float az) {
                    float k2 = 0.0f;
                    for (int a = 0; a < 3; ++a) {
                        const float kv = ax * rf[a] + ay * rf[3 + a] + az * rf[6 + a];
                        k2 += kv * kv;
                    }
                    if (disp) {
                        const float kk = __builtin_sqrtf(k2);
                        float ec, ex;
                        Math<float>::erfc_exp((float)pi_a * kk, ec, ex);
                        return (float)dpref * ((float)dc1 * ec * kk * k2 + ex * ((float)da3 - (float)dc2 * k2)) /
                               mod;
                    }
                    return pref_f * __builtin_amdgcn_exp2f(-pi2_a2_log2e * k2) / (k2 * mod);
                };
                float infl = influence(mx, my, mz);
                if (nyq) infl = 0.5f * (infl + influence(-sx * mx, -sy * my, -sz * mz));
                S.x = (R)((float)S.x * infl);
                S.y = (R)((float)S.y * infl);
                continue;
            }
            const double mx = kx < (nx + 1) / 2 ? kx : kx - nx;
            const double my = y < (ny + 1) / 2 ? y : y - ny;
            const double mz = kz < (nz + 1) / 2 ? kz : kz - nz;
            const double mod = (double)mod_x[kx] * (double)mod_y[y] * (double)mod_z[kz];
            auto influence = [&](double ax, double ay, double az) {
                double k2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    double kv = ax * b.recip[0 * 3 + a] + ay * b.recip[1 * 3 + a] + az * b.recip[2 * 3 + a];
                    k2 += kv * kv;
                }
                if (disp) {
                    const double kk = sqrt(k2), bb = pi_a * kk;
                    return dpref / mod * (dc1 * erfc(bb) * kk * k2 + exp(-bb * bb) * (da3 - dc2 * k2));
                }
                return pref * exp(-pi2_a2 * k2) / (k2 * mod);
            };
            double infl = influence(mx, my, mz);
            if (nyq) infl = 0.5 * (infl + influence(-sx * mx, -sy * my, -sz * mz));
            const double wz = (kz == 0 || 2 * kz == nz) ? 1.0 : 2.0;
            const double s2 = (double)S.x * S.x + (double)S.y * S.y;
            eacc += wz * 0.5 * infl * s2;
            S.x = (R)(S.x * infl);
            S.y = (R)(S.y * infl);
        }
        if (!F32) {
            double et = wave_sum(eacc);
            if ((threadIdx.x & 63) == 0) acc_add(energy + si * E_SLOTS + E_RECIP, et, ENERGY_SCALE, status);
        }
        __syncthreads();
        Cx<R>* other = res == bufA ? bufB : bufA;
        res = fft_lds<R>(res, other, n, rad, roots, true, T);
    }

    // Store.
    if (MODE == MODE_Z_R2C) {
        for (int e = threadIdx.x; e < LINES * nzh; e += FFT_THREADS) {
            const int li = e / nzh, el = e - li * nzh;
            const int L = line0 + li;
            if (L >= nlines) continue;
            const int x = L / nyp, y = 2 * (L - x * nyp);
            Cx<R> a, b;
            unpack_pair(res[li * n + el], res[li * n + (el ? n - el : 0)], a, b);
            Cx<R>* row = cgrid + ((size_t)x * ny + y) * nzh;
            row[el] = a;
            if (y + 1 < ny) row[nzh + el] = b;
        }
        return;
    }
    for (int e = threadIdx.x; e < LINES * n; e += FFT_THREADS) {
        int li, el;
        if (zmode) {
            li = e / n;
            el = e - li * n;
        } else {
            li = e % LINES;
            el = e / LINES;
        }
        const int L = line0 + li;
        if (L >= nlines) continue;
        const Cx<R> v = res[li * n + el];
        if (MODE == MODE_Z_C2R) {
            const int x = L / nyp, y = 2 * (L - x * nyp);
            const size_t row = (size_t)x * ny + y;
            phi[row * nz + el] = v.x;
            if (y + 1 < ny) phi[(row + 1) * nz + el] = v.y;
        } else if (ymode) {
            const int x = L / nzh, kz = L - x * nzh;
            cgrid[((size_t)x * ny + el) * nzh + kz] = v;
        } else {
            const int y = L / nzh, kz = L - y * nzh;
            cgrid[((size_t)el * ny + y) * nzh + kz] = v;
        }
    }
}

// Plane passes: one workgroup per x plane of a system transforms the plane
// along z and along y in LDS.  Forward (DIR 0): pairs of real z lines to their
// half spectra (kept as P[kz][y]), then y lines, stored as cgrid[x][ky][kz].
// Inverse (DIR 1): y lines of the half spectrum, then pairs of z lines rebuilt
// from it by Hermitian symmetry and transformed to real values.  Every line
// goes through fft_lds (and the pair packing) with the data the corresponding
// line pass would load, so the results are the same bits as the five-pass
// form.
//
// LDS (complex elements): roots along z [nz] and y [ny], P and Q
// [lyp x ny] (lyp = nzh rounded up to a multiple of the y lines per call),
// and the ping-pong buffers of lz z-line pairs [2 x lz x nz], which share
// the one of P and Q that does not hold live data when they fit in it.
constexpr int PLANE_THREADS = 1024;
constexpr size_t PLANE_LDS_MAX = 160 * 1024 - 256;  // gfx950 LDS per workgroup, less the static arrays

struct PlaneLayout {
    int lyp;
    size_t pq;      // elements of P (and of Q)
    bool z_shared;  // the z buffers fit in the free one of P and Q
    size_t elems;   // total complex elements
};

__host__ __device__ inline PlaneLayout plane_layout(int ny, int nz, int nzh, int lz, int ly) {
    PlaneLayout L;
    L.lyp = (nzh + ly - 1) / ly * ly;
    L.pq = (size_t)L.lyp * ny;
    const size_t zb = 2 * (size_t)lz * nz;
    L.z_shared = zb <= L.pq;
    L.elems = (size_t)nz + ny + 2 * L.pq + (L.z_shared ? 0 : zb);
    return L;
}

template <class R, class G, int DIR>
__global__ __launch_bounds__(PLANE_THREADS) void k_fft_plane(const int* __restrict__ plane_off, int nsys,
                                                            const SysInfo* __restrict__ sys, G* __restrict__ qgrid,
                                                            Cx<R>* cgrid, R* phi,
                                                            const Cx<R>* __restrict__ roots_all, int* clear,
                                                            int nclear) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ int s_rz[MAX_STAGES], s_ry[MAX_STAGES];
    // The forward pass also clears the spread's bin counts (every spread
    // workgroup is done by now), as the first line pass does.
    if (DIR == 0 && clear)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nclear; i += gridDim.x * blockDim.x) clear[i] = 0;
    const int si = find_sys(plane_off, nsys, blockIdx.x);
    const SysInfo& S = sys[si];
    const int ny = S.grid[1], nz = S.grid[2], nzh = S.nzh;
    const int x = blockIdx.x - plane_off[si];
    constexpr int prec = sizeof(R) == 8 ? 1 : 0;
    const int lz = S.plane_lines[prec][0], ly = S.plane_lines[prec][1];
    const int nyp = (ny + 1) / 2;  // z line pairs
    const PlaneLayout L = plane_layout(ny, nz, nzh, lz, ly);
    Cx<R>* rz = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* ry = rz + nz;
    Cx<R>* P = ry + ny;
    Cx<R>* Q = P + L.pq;
    Cx<R>* extra = Q + L.pq;
    if (threadIdx.x < MAX_STAGES) {
        s_rz[threadIdx.x] = S.radix[2][threadIdx.x];
        s_ry[threadIdx.x] = S.radix[1][threadIdx.x];
    }
    for (int i = threadIdx.x; i < nz; i += PLANE_THREADS) rz[i] = roots_all[S.roots_off[2] + i];
    for (int i = threadIdx.x; i < ny; i += PLANE_THREADS) ry[i] = roots_all[S.roots_off[1] + i];
    qgrid += S.qgrid_off + (size_t)x * ny * nz;
    phi += S.qgrid_off + (size_t)x * ny * nz;
    cgrid += S.cgrid_off + (size_t)x * ny * nzh;
    const int Tz = PLANE_THREADS / lz, Ty = PLANE_THREADS / ly;

    if (DIR == 0) {
        Cx<R>* Z = L.z_shared ? Q : extra;
        __syncthreads();  // roots and radices
        for (int p0 = 0; p0 < nyp; p0 += lz) {
            // Line pairs (y, y + 1) = (2p, 2p + 1), as the line passes pair them.
            for (int e = threadIdx.x; e < lz * nz; e += PLANE_THREADS) {
                const int li = e / nz, el = e - li * nz;
                const int y = 2 * (p0 + li);
                Cx<R> v = {0, 0};
                if (y < ny) {
                    G& g0 = qgrid[(size_t)y * nz + el];  // cleared as read
                    v.x = (R)grid_real(g0);
                    g0 = (G)0;
                    if (y + 1 < ny) {
                        G& g1 = qgrid[(size_t)(y + 1) * nz + el];
                        v.y = (R)grid_real(g1);
                        g1 = (G)0;
                    }
                }
                Z[e] = v;
            }
            __syncthreads();
            const Cx<R>* res = fft_lds<R>(Z, Z + (size_t)lz * nz, nz, s_rz, rz, false, Tz);
            for (int e = threadIdx.x; e < lz * nzh; e += PLANE_THREADS) {
                const int li = e / nzh, kz = e - li * nzh;
                const int y = 2 * (p0 + li);
                if (y < ny) {
                    Cx<R> a, b;
                    unpack_pair(res[(size_t)li * nz + kz], res[(size_t)li * nz + (kz ? nz - kz : 0)], a, b);
                    P[(size_t)kz * ny + y] = a;
                    if (y + 1 < ny) P[(size_t)kz * ny + y + 1] = b;
                }
            }
            __syncthreads();
        }
        const Cx<R>* res = P;
        for (int k0 = 0; k0 < L.lyp; k0 += ly)
            res = fft_lds<R>(P + (size_t)k0 * ny, Q + (size_t)k0 * ny, ny, s_ry, ry, false, Ty) - (size_t)k0 * ny;
        for (int e = threadIdx.x; e < ny * nzh; e += PLANE_THREADS) {
            const int ky = e / nzh, kz = e - ky * nzh;
            cgrid[e] = res[(size_t)kz * ny + ky];
        }
    } else {
        for (int e = threadIdx.x; e < ny * nzh; e += PLANE_THREADS) {
            const int ky = e / nzh, kz = e - ky * nzh;
            P[(size_t)kz * ny + ky] = cgrid[e];
        }
        __syncthreads();
        const Cx<R>* res = P;
        for (int k0 = 0; k0 < L.lyp; k0 += ly)
            res = fft_lds<R>(P + (size_t)k0 * ny, Q + (size_t)k0 * ny, ny, s_ry, ry, true, Ty) - (size_t)k0 * ny;
        Cx<R>* Z = L.z_shared ? (res == P ? Q : P) : extra;
        // The half spectrum of line y at kz is res[kz][y].
        auto half_at = [&](int y, int el) {
            if (el < nzh) return res[(size_t)el * ny + y];
            const Cx<R> v = res[(size_t)(nz - el) * ny + y];
            return Cx<R>{v.x, -v.y};
        };
        for (int p0 = 0; p0 < nyp; p0 += lz) {
            for (int e = threadIdx.x; e < lz * nz; e += PLANE_THREADS) {
                const int li = e / nz, el = e - li * nz;
                const int y = 2 * (p0 + li);
                Cx<R> v = {0, 0};
                if (y < ny) v = pack_pair(half_at(y, el), y + 1 < ny ? half_at(y + 1, el) : Cx<R>{0, 0});
                Z[e] = v;
            }
            __syncthreads();
            const Cx<R>* out = fft_lds<R>(Z, Z + (size_t)lz * nz, nz, s_rz, rz, true, Tz);
            for (int e = threadIdx.x; e < lz * nz; e += PLANE_THREADS) {
                const int li = e / nz, el = e - li * nz;
                const int y = 2 * (p0 + li);
                if (y < ny) {
                    phi[(size_t)y * nz + el] = out[e].x;
                    if (y + 1 < ny) phi[(size_t)(y + 1) * nz + el] = out[e].y;
                }
            }
            __syncthreads();
        }
    }
}

std::vector<int> factorize(int n) {
    std::vector<int> r;
    while (n % 8 == 0) {
        r.push_back(8);
        n /= 8;
    }
    while (n % 4 == 0) {
        r.push_back(4);
        n /= 4;
    }
    while (n % 2 == 0) {
        r.push_back(2);
        n /= 2;
    }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    if (n > 1) r.push_back(n);
    return r;
}

// Most lines per workgroup whose two LDS buffers fit in 60 KB (at most 16).
int max_lines(int n, size_t csize) {
    int lines = 16;
    while (lines > 1 && (2 * (size_t)lines * n + n) * csize > 60000) lines >>= 1;
    return lines;
}

template <class R, class G, class A, int MODE>
void launch_fft(const PmeDev& pm, const SysInfo* sys, G* qgrid, Cx<R>* cgrid, R* phi, const R* roots, const R* moduli,
                A* energy, unsigned* status, hipStream_t s, int* clear = nullptr, int nclear = 0) {
    const int prec = sizeof(R) == 8 ? 1 : 0;
    const bool zmode = MODE == MODE_Z_R2C || MODE == MODE_Z_C2R;
    const bool ymode = MODE == MODE_Y_FWD || MODE == MODE_Y_INV;
    const int axis = zmode ? 2 : (ymode ? 1 : 0);
    hipLaunchKernelGGL((k_fft<R, G, A, MODE>), dim3(pm.fft_blocks[prec][MODE]), dim3(FFT_THREADS),
                       pm.fft_shmem[prec][axis], s, pm.fft_off[prec][MODE].p, pm.nsys, sys, qgrid, cgrid, phi,
                       reinterpret_cast<const Cx<R>*>(roots), moduli, energy, status, clear, nclear);
}

template <class R, class G, int DIR>
void launch_plane(const PmeDev& pm, const SysInfo* sys, G* qgrid, Cx<R>* cgrid, R* phi, const R* roots,
                  hipStream_t s, int* clear = nullptr, int nclear = 0) {
    constexpr int prec = sizeof(R) == 8 ? 1 : 0;
    hipLaunchKernelGGL((k_fft_plane<R, G, DIR>), dim3(pm.plane_blocks), dim3(PLANE_THREADS), pm.plane_shmem[prec], s,
                       pm.plane_off.p, pm.nsys, sys, qgrid, cgrid, phi, reinterpret_cast<const Cx<R>*>(roots), clear,
                       nclear);
}

template <class R>
R* pme_weights(PmeDev& pm);
template <>
float* pme_weights<float>(PmeDev& pm) {
    return pm.st_wf.p;
}
template <>
double* pme_weights<double>(PmeDev& pm) {
    return pm.st_wd.p;
}

template <class R, class G, class A>
void run_t(PmeDev& pm, const NlistDev& nl, bool atom_order, const double* pos, const double* charge,
           const SysInfo* sys, G* qgrid,
           R* cgrid, R* phi, const R* roots, const R* moduli, A* force, A* energy, unsigned* status, int phases,
           hipStream_t s) {
    // Bin counts and the charge grid are zero between runs: set so at
    // allocation, and cleared by the first FFT pass after each spread (the
    // grid point by point as the pass reads it), or by fills when a run
    // spreads without transforming.
    const bool own_order = pme_order_ready(pm, nl);
    const int* own = own_order ? pm.order.p + (size_t)pm.order_sel * nl.nslots : nullptr;
    if (phases & PME_SPREAD) {
        const int* order = atom_order ? own : nl.sorted_atom.p;
        const int small = pm.max_charge <= SPREAD32_MAX_CHARGE ? 1 : 0;
        // The atom-owned spread adds to the grid with device atomics, so it
        // serves the fixed-point grids only (FP64 atomic sums would depend on
        // their order).
        constexpr bool fixed_grid = sizeof(G) == sizeof(long long) && G(0.5) == G(0);
        if (pm.spread_atoms && fixed_grid) {
            const int blocks = order ? nl.nchunks : (pm.natoms + PME_CHUNK - 1) / PME_CHUNK;
            if (blocks > 0)
                hipLaunchKernelGGL((k_spread_atoms<R, G>), dim3(blocks), dim3(ATOM_SPREAD_THREADS), 0, s,
                                   order ? nl.chunk_off.p : nullptr, pm.nsys, order, pm.natoms, pos, charge,
                                   pm.atom_sys, sys, qgrid, small, status);
        } else {
            R* st_w = pme_weights<R>(pm);
            const int nitems = order ? nl.nslots : pm.natoms;
            hipLaunchKernelGGL((k_pme_bin<R>), dim3((nitems + 255) / 256), dim3(256), 0, s, nitems, order, pos,
                               charge, pm.atom_sys, sys, pm.st_idx.p, st_w, pm.bin_count.p, pm.bins.p, pm.bin_cap,
                               status);
            hipLaunchKernelGGL((k_spread_tiles<R, G>), dim3(pm.ntiles), dim3(SPREAD_THREADS), 0, s, pm.tile_off.p,
                               pm.nsys, sys, pm.st_idx.p, st_w, charge, pm.bin_count.p, pm.bins.p, pm.bin_cap,
                               qgrid, small, status);
            if (!(phases & PME_FFT)) fill_async<int>(pm.bin_count.p, (size_t)pm.ntiles, 0, s);
        }
        if (!(phases & PME_FFT)) fill_async<G>(qgrid, (size_t)pm.qpoints, (G)0, s);
    }
    if (phases & PME_FFT) {
        Cx<R>* cg = reinterpret_cast<Cx<R>*>(cgrid);
        int* clear = (phases & PME_SPREAD) ? pm.bin_count.p : nullptr;
        constexpr int prec = sizeof(R) == 8 ? 1 : 0;
        if (pm.plane[prec]) {
            launch_plane<R, G, 0>(pm, sys, qgrid, cg, phi, roots, s, clear, pm.ntiles);
            launch_fft<R, G, A, MODE_X_CONV>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s);
            launch_plane<R, G, 1>(pm, sys, qgrid, cg, phi, roots, s);
        } else {
            launch_fft<R, G, A, MODE_Z_R2C>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s, clear,
                                            pm.ntiles);
            launch_fft<R, G, A, MODE_Y_FWD>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s);
            launch_fft<R, G, A, MODE_X_CONV>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s);
            launch_fft<R, G, A, MODE_Y_INV>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s);
            launch_fft<R, G, A, MODE_Z_C2R>(pm, sys, qgrid, cg, phi, roots, moduli, energy, status, s);
        }
    }
    if (phases & PME_GATHER) {
        // With atom_order the gather follows the same order as the spread:
        // it may then run while a list build rewrites nl.sorted_atom.
        const int* order = atom_order && own_order ? own : nl.sorted_atom.p;
        hipLaunchKernelGGL((k_gather5<R, A>), dim3(nl.nchunks), dim3(GATHER_ATOMS * 5), 0, s, nl.chunk_off.p, pm.nsys,
                           order, pos, charge, sys, phi, force, status);
    }
    OMM_HIP(hipGetLastError());
}

// The mesh potential at the stencil of each listed atom of system 0,
// sum_p phi_p w_x w_y w_z (PmeCache::charge_derivatives, pme.rs:117-137: the
// derivative of the mesh energy with respect to that atom's charge), from the
// FP64 potential of the last FP64 transform.  One thread per listed atom.
__global__ void k_mesh_potential(int count, const int* __restrict__ atoms, const double* __restrict__ pos,
                                 const SysInfo* __restrict__ sys, const double* __restrict__ phi, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const SysInfo& S = sys[0];
    const int nx = S.grid[0], ny = S.grid[1], nz = S.grid[2];
    int base[3];
    double w[3][5], dw[3][5];
    stencil<double>(pos + 3 * (size_t)atoms[k], S.box, S.grid, base, w, dw);
    const double* g = phi + S.qgrid_off;
    double sum = 0.0;
    for (int ix = 0; ix < 5; ++ix) {
        const int gx = (base[0] + ix) % nx;
        for (int iy = 0; iy < 5; ++iy) {
            const int gy = (base[1] + iy) % ny;
            const double* row = g + ((size_t)gx * ny + gy) * nz;
            for (int iz = 0; iz < 5; ++iz) sum += row[(base[2] + iz) % nz] * w[0][ix] * w[1][iy] * w[2][iz];
        }
    }
    out[k] = sum;
}

}  // namespace

void pme_potential_at(const PmeDev& pm, const SysInfo* sys_dev, const double* pos, const int* atoms, int count,
                      double* out, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_mesh_potential, dim3((count + 255) / 256), dim3(256), 0, s, count, atoms, pos, sys_dev,
                       pm.phi_d.p, out);
    OMM_HIP(hipGetLastError());
}

std::vector<double> bspline_moduli(int n) {
    // Order-5 spline at integer points: M5(1..4) = 1/24, 11/24, 11/24, 1/24
    // (the weights of a zero fractional coordinate).
    const double w[5] = {1.0 / 24.0, 11.0 / 24.0, 11.0 / 24.0, 1.0 / 24.0, 0.0};
    std::vector<double> values(n);
    for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        for (int i = 0; i < 5; ++i) {
            double t = 2.0 * PI * (double)k * (double)i / (double)n;
            re += w[i] * std::cos(t);
            im += w[i] * std::sin(t);
        }
        values[k] = re * re + im * im;
    }
    for (int k = 0; k < n; ++k)
        if (values[k] < 1e-7) values[k] = 0.5 * (values[(k + n - 1) % n] + values[(k + 1) % n]);
    return values;
}

void pm_set_bin_cap(PmeDev& pm, int cap) {
    pm.bin_cap = cap;
    pm.bins.alloc((size_t)pm.ntiles * cap);
}

void pme_grow_bins(PmeDev& pm) { pm_set_bin_cap(pm, 2 * pm.bin_cap); }

namespace {
// Both halves of each given order.
__global__ void k_copy_order(int n, const int* __restrict__ src, int* a, int* b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int v = src[k];
    if (a) a[k] = a[n + k] = v;
    if (b) b[k] = b[n + k] = v;
}
}  // namespace

void pme_follow_order(PmeDev& pm, PmeDev* other, const NlistDev& nl, hipStream_t s) {
    const size_t n2 = 2 * (size_t)nl.nslots;
    int* a = nl.nslots > 0 && pm.order.n == n2 ? pm.order.p : nullptr;
    int* b = nl.nslots > 0 && other && other->order.n == n2 ? other->order.p : nullptr;
    if (!a && !b) return;
    hipLaunchKernelGGL(k_copy_order, dim3((nl.nslots + 255) / 256), dim3(256), 0, s, nl.nslots, nl.sorted_atom.p, a,
                       b);
    OMM_HIP(hipGetLastError());
    if (a) pm.order_gen = nl.generation;
    if (b) other->order_gen = nl.generation;
}

bool pme_order_ready(const PmeDev& pm, const NlistDev& nl) {
    return nl.nslots > 0 && pm.order_gen == nl.generation && pm.order.n == 2 * (size_t)nl.nslots;
}

int* pme_next_order(PmeDev& pm, const NlistDev& nl) {
    pm.order_sel ^= 1;
    return pm.order.p + (size_t)pm.order_sel * nl.nslots;
}

void pme_take_order(PmeDev& pm, const NlistDev& nl, hipStream_t s) {
    if (nl.nslots == 0 || nl.sorted_atom.n < (size_t)nl.nslots) return;
    if (pm.order.n != 2 * (size_t)nl.nslots) pm.order.alloc(2 * (size_t)nl.nslots);
    pm.order_gen = -1;  // pme_follow_order sets it
    pme_follow_order(pm, nullptr, nl, s);
}

namespace {
const int PASS_AXIS[5] = {2, 1, 0, 1, 2};  // MODE_Z_R2C, Y_FWD, X_CONV, Y_INV, Z_C2R

int fft_nlines(const SysInfo& S, int a) {
    // z passes: pairs of real lines, (ny + 1) / 2 per x plane.
    return a == 0 ? S.grid[1] * S.nzh : (a == 1 ? S.grid[0] * S.nzh : S.grid[0] * ((S.grid[1] + 1) / 2));
}
}  // namespace

int pme_set_fft_lines(PmeDev& pm, std::vector<SysInfo>& sys, int prec, int cap, hipStream_t s) {
    const size_t csize = prec ? 2 * sizeof(double) : 2 * sizeof(float);
    if (cap <= 0) {
        // The most lines that fit in LDS, lowered for the whole batch until
        // every pass has ~2048 workgroups (8 per CU) or two lines each.
        for (cap = 16;; cap >>= 1) {
            long long fewest = LLONG_MAX;
            for (int k = 0; k < 5; ++k) {
                long long blocks = 0;
                for (const SysInfo& S : sys) {
                    const int a = PASS_AXIS[k];
                    const int l = std::min(cap, max_lines(S.grid[a], csize));
                    blocks += (fft_nlines(S, a) + l - 1) / l;
                }
                fewest = std::min(fewest, blocks);
            }
            if (fewest >= 2048 || cap <= 2) break;
        }
    }
    std::vector<int> off[5];
    for (int k = 0; k < 5; ++k) off[k].assign(1, 0);
    for (int a = 0; a < 3; ++a) pm.fft_shmem[prec][a] = 0;
    for (SysInfo& S : sys) {
        for (int a = 0; a < 3; ++a) {
            S.fft_lines[prec][a] = std::min(cap, max_lines(S.grid[a], csize));
            pm.fft_shmem[prec][a] =
                std::max(pm.fft_shmem[prec][a], (2 * (size_t)S.fft_lines[prec][a] * S.grid[a] + S.grid[a]) * csize);
        }
        for (int k = 0; k < 5; ++k) {
            const int a = PASS_AXIS[k];
            off[k].push_back(off[k].back() + (fft_nlines(S, a) + S.fft_lines[prec][a] - 1) / S.fft_lines[prec][a]);
        }
    }
    for (int k = 0; k < 5; ++k) {
        pm.fft_off[prec][k].upload(off[k], s);
        pm.fft_blocks[prec][k] = off[k].back();
    }
    pm.fft_cap[prec] = cap;
    return cap;
}

bool pme_set_fft_plane(PmeDev& pm, std::vector<SysInfo>& sys, int prec, int use) {
    const size_t csize = prec ? 2 * sizeof(double) : 2 * sizeof(float);
    size_t shmem = 0;
    bool ok = use != 0 && !sys.empty();
    for (SysInfo& S : sys) {
        if (!ok) break;
        const int ny = S.grid[1], nz = S.grid[2], nzh = S.nzh;
        // y lines per call: all of the half spectrum's lines when the threads
        // of a line stay at most ny, fewer calls otherwise; z line pairs per call:
        // as many as fit, down to the fewest that keep the threads of a line
        // at most nz.
        auto pow2_at_least = [](int v) {
            int p = 1;
            while (p < v) p <<= 1;
            return p;
        };
        const int ly = std::min(PLANE_THREADS, std::max(pow2_at_least(nzh), pow2_at_least((PLANE_THREADS + ny - 1) / ny)));
        const int lz_min = std::min(PLANE_THREADS, pow2_at_least((PLANE_THREADS + nz - 1) / nz));
        int lz = std::min(PLANE_THREADS, std::max(lz_min, pow2_at_least((ny + 1) / 2)));
        while (lz > lz_min && plane_layout(ny, nz, nzh, lz, ly).elems * csize > PLANE_LDS_MAX) lz >>= 1;
        const size_t bytes = plane_layout(ny, nz, nzh, lz, ly).elems * csize;
        if (bytes > PLANE_LDS_MAX) {
            ok = false;
            break;
        }
        S.plane_lines[prec][0] = lz;
        S.plane_lines[prec][1] = ly;
        shmem = std::max(shmem, bytes);
    }
    pm.plane[prec] = ok;
    pm.plane_shmem[prec] = ok ? shmem : 0;
    return ok;
}

int pme_set_spread_edge(PmeDev& pm, std::vector<SysInfo>& sys, int edge, hipStream_t s) {
    if (edge <= 0) {
        edge = 8;
        for (int e : {8, 6}) {
            long long total = 0;
            for (const SysInfo& S : sys)
                total += (long long)tiles_for(S.grid[0], e) * tiles_for(S.grid[1], e) * tiles_for(S.grid[2], e);
            edge = e;
            if (total >= SPREAD_MIN_TILES) break;
        }
    }
    std::vector<int> toff{0};
    int bin_cap = 64;
    for (SysInfo& S : sys) {
        S.ptile = edge;
        for (int a = 0; a < 3; ++a) S.tiles[a] = tiles_for(S.grid[a], edge);
        S.tile_off = toff.back();
        const int nt = S.tiles[0] * S.tiles[1] * S.tiles[2];
        toff.push_back(toff.back() + nt);
        // Twice the atoms of the largest tile at the system's mean density
        // (the last tile of an axis is up to twice as wide as the others).
        double largest = 1.0;
        for (int a = 0; a < 3; ++a) {
            const int last = S.grid[a] - (S.tiles[a] - 1) * edge;
            largest *= (double)std::max(last, std::min(edge, S.grid[a])) / (double)S.grid[a];
        }
        bin_cap = std::max(bin_cap, (int)(2.0 * largest * S.n) + 64);
    }
    pm.ntiles = toff.back();
    pm.tile_off.upload(toff, s);
    pm.bin_count.alloc((size_t)pm.ntiles);
    fill_async<int>(pm.bin_count.p, (size_t)pm.ntiles, 0, s);
    pm_set_bin_cap(pm, bin_cap);
    pm.spread_edge = edge;
    return edge;
}

void pme_setup(PmeDev& pm, std::vector<SysInfo>& sys, int natoms, const int* atom_sys, bool mixed, hipStream_t s) {
    const int nsys = (int)sys.size();
    pm.nsys = nsys;
    pm.natoms = natoms;
    pm.atom_sys = atom_sys;
    std::vector<double> roots, moduli;
    long long qpts = 0, cpts = 0;
    for (int si = 0; si < nsys; ++si) {
        SysInfo& S = sys[si];
        for (int a = 0; a < 3; ++a)
            if (S.grid[a] < 5 || S.grid[a] > 1024)
                throw Error(OMM_ERR_INVALID, "PME grid dimensions must be in [5, 1024]");
        S.nzh = S.grid[2] / 2 + 1;
        S.qgrid_off = qpts;
        S.cgrid_off = cpts;
        qpts += (long long)S.grid[0] * S.grid[1] * S.grid[2];
        cpts += (long long)S.grid[0] * S.grid[1] * S.nzh;
        for (int a = 0; a < 3; ++a) {
            const int n = S.grid[a];
            std::vector<int> f = factorize(n);
            if ((int)f.size() >= MAX_STAGES) throw Error(OMM_ERR_INVALID, "PME grid dimension has too many factors");
            for (int k = 0; k < MAX_STAGES; ++k) S.radix[a][k] = k < (int)f.size() ? f[k] : 0;
            S.roots_off[a] = (int)(roots.size() / 2);
            for (int m = 0; m < n; ++m) {
                double t = -2.0 * PI * (double)m / (double)n;
                roots.push_back(std::cos(t));
                roots.push_back(std::sin(t));
            }
            S.moduli_off[a] = (int)moduli.size();
            auto mod = bspline_moduli(n);
            moduli.insert(moduli.end(), mod.begin(), mod.end());
        }
    }
    pm.qpoints = qpts;
    pm.cpoints = cpts;
    // Default grouping: the largest line cap with ~2048 workgroups per FFT
    // pass, and the largest spread tile edge with SPREAD_MIN_TILES tiles.
    pme_set_fft_lines(pm, sys, 0, 0, s);
    pme_set_fft_lines(pm, sys, 1, 0, s);
    {
        std::vector<int> poff{0};
        for (const SysInfo& S : sys) poff.push_back(poff.back() + S.grid[0]);
        pm.plane_off.upload(poff, s);
        pm.plane_blocks = poff.back();
        // Dynamic LDS above 64 KB per workgroup.
        const int lds = (int)PLANE_LDS_MAX;
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<float, long long, 0>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<float, long long, 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<double, long long, 0>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<double, long long, 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<double, double, 0>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        OMM_HIP(hipFuncSetAttribute((const void*)&k_fft_plane<double, double, 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        for (int prec = 0; prec < 2; ++prec) pme_set_fft_plane(pm, sys, prec, -1);
    }
    pme_set_spread_edge(pm, sys, 0, s);
    pm.st_idx.alloc(4 * (size_t)pm.natoms);
    if (mixed) pm.st_wf.alloc(16 * (size_t)pm.natoms);
    pm.st_wd.alloc(16 * (size_t)pm.natoms);
    pm.roots_d.upload(roots, s);
    pm.moduli_d.upload(moduli, s);
    pm.cgrid_d.alloc(2 * (size_t)cpts);
    pm.phi_d.alloc((size_t)qpts);
    if (mixed) {
        std::vector<float> rf(roots.begin(), roots.end()), mf(moduli.begin(), moduli.end());
        pm.roots_f.upload(rf, s);
        pm.moduli_f.upload(mf, s);
        pm.qgrid_fx.alloc((size_t)qpts);
        fill_async<long long>(pm.qgrid_fx.p, (size_t)qpts, 0ll, s);
        pm.cgrid_f.alloc(2 * (size_t)cpts);
        pm.phi_f.alloc((size_t)qpts);
    } else {
        pm.qgrid_d.alloc((size_t)qpts);
        fill_async<double>(pm.qgrid_d.p, (size_t)qpts, 0.0, s);
    }
    OMM_HIP(hipStreamSynchronize(s));
}

void pme_run(PmeDev& pm, int mode, const NlistDev& nl, const double* pos, const double* charge,
             const SysInfo* sys_dev, long long* force_fx, double* force_d, long long* energy_fx, double* energy_d,
             unsigned* status, hipStream_t s, int phases, bool atom_order) {
    if (mode == MESH_F32)
        run_t<float, long long, long long>(pm, nl, atom_order, pos, charge, sys_dev, pm.qgrid_fx.p, pm.cgrid_f.p, pm.phi_f.p,
                                           pm.roots_f.p, pm.moduli_f.p, force_fx, energy_fx, status, phases, s);
    else if (mode == MESH_F64_FIXED)
        run_t<double, long long, long long>(pm, nl, atom_order, pos, charge, sys_dev, pm.qgrid_fx.p, pm.cgrid_d.p, pm.phi_d.p,
                                            pm.roots_d.p, pm.moduli_d.p, force_fx, energy_fx, status, phases, s);
    else
        run_t<double, double, double>(pm, nl, atom_order, pos, charge, sys_dev, pm.qgrid_d.p, pm.cgrid_d.p, pm.phi_d.p,
                                      pm.roots_d.p, pm.moduli_d.p, force_d, energy_d, status, phases, s);
}

}  // namespace ommhip
