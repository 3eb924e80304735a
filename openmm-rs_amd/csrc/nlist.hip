// Tile neighbour list: deterministic spatial sort, block bounds, ballot-compacted
// j-lists and the task prefix sum of the pair kernel.
#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>

#include "nlist.h"

namespace ommhip {

namespace {


__device__ __forceinline__ void wrap_fast(double d[3], const Box& b) { wrap_nearest(d, b); }

// Squared distance between the periodic images of an axis-aligned box (centre
// offset dc, half-extents h) and the origin; conservative for triclinic boxes
// by scanning the 27 images around the reduced nearest image.
__device__ __forceinline__ double box_gap2(const double dc_in[3], const double h[3], const Box& b) {
    double dc[3] = {dc_in[0], dc_in[1], dc_in[2]};
    wrap_fast(dc, b);
    if (!b.triclinic) {
        double g2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double g = fabs(dc[a]) - h[a];
            if (g > 0.0) g2 += g * g;
        }
        return g2;
    }
    double best = 1e300;
    for (int nx = -1; nx <= 1; ++nx)
        for (int ny = -1; ny <= 1; ++ny)
            for (int nz = -1; nz <= 1; ++nz) {
                double g2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    double v = dc[a] + nx * b.v[0 * 3 + a] + ny * b.v[1 * 3 + a] + nz * b.v[2 * 3 + a];
                    double g = fabs(v) - h[a];
                    if (g > 0.0) g2 += g * g;
                }
                best = fmin(best, g2);
            }
    return best;
}

// flag[0]: some system needs a rebuild; flag[1 + s]: system s does.  force:
// flag every system.
__global__ void k_check(const double* __restrict__ pos, const double* __restrict__ pos_build,
                        const int* __restrict__ atom_sys, int n, int nsys, double lim2, int force, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    int si = 0;
    if (i < n) {
        double d2 = skin_d2(pos[3 * i] - pos_build[3 * i], pos[3 * i + 1] - pos_build[3 * i + 1],
                            pos[3 * i + 2] - pos_build[3 * i + 2]);
        moved = !(d2 <= lim2);  // NaN also requests a rebuild
        si = atom_sys[i];
    }
    wave_raise_flags(flag, moved, si);
    if (force && i < nsys) {
        atomicOr(flag + 1 + i, 1);
        atomicOr(flag, 1);
    }
}

__device__ __forceinline__ void unit_key_of(int u, const double* __restrict__ pos, const int* __restrict__ unit_ptr,
                                            const int* __restrict__ unit_atoms, const int* __restrict__ unit_sys,
                                            const SysInfo* __restrict__ sys, const int* __restrict__ cell_rank,
                                            int* unit_key, int* cell_count, const int* __restrict__ flag) {
    const int si = unit_sys[u];
    if (!flag[1 + si]) {
        // Not rebuilt: the previous key reproduces the previous order.
        atomicAdd(&cell_count[unit_key[u]], 1);
        return;
    }
    int a = unit_atoms[unit_ptr[u]];
    const SysInfo& S = sys[si];
    const Box& b = S.box;
    int c[3];
    for (int ax = 0; ax < 3; ++ax) {
        double f = pos[3 * a] * b.recip[ax * 3] + pos[3 * a + 1] * b.recip[ax * 3 + 1] +
                   pos[3 * a + 2] * b.recip[ax * 3 + 2];
        f -= floor(f);
        const int nc = S.ncell[ax];
        int v = (int)(f * nc);
        c[ax] = v < 0 ? 0 : (v >= nc ? nc - 1 : v);
    }
    int key = S.cell_off + cell_rank[S.cell_off + (c[0] * S.ncell[1] + c[1]) * S.ncell[2] + c[2]];
    unit_key[u] = key;
    atomicAdd(&cell_count[key], 1);
}

__global__ void k_unit_keys(const double* __restrict__ pos, const int* __restrict__ unit_ptr,
                            const int* __restrict__ unit_atoms, const int* __restrict__ unit_sys, int nunits,
                            const SysInfo* __restrict__ sys, const int* __restrict__ cell_rank, int* unit_key,
                            int* cell_count, const int* __restrict__ flag) {
    if (!flag[0]) return;
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    unit_key_of(u, pos, unit_ptr, unit_atoms, unit_sys, sys, cell_rank, unit_key, cell_count, flag);
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total, in three
// launches: per-chunk sums, a one-workgroup scan of the chunk sums, and the
// per-chunk scans with their offsets.  copy (optional) receives the same
// values; zero_input clears the input after reading it (the cell counts start
// the next build at zero without a fill kernel).
constexpr int SCAN_WG = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_CHUNK = SCAN_WG * SCAN_PER;

// Inclusive scan of v over the workgroup (SCAN_WG threads); returns it and
// leaves the workgroup total in *total.
__device__ __forceinline__ int block_scan_inclusive(int v, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < SCAN_WG / 64; ++w) {
        const int x = s_wave[w];
        if (w < wave) before += x;
        all += x;
    }
    *total = all;
    return v + before;
}

__global__ __launch_bounds__(SCAN_WG) void k_scan_reduce(const int* __restrict__ in, int n, int* partial,
                                                         const int* flag) {
    if (!flag[0]) return;
    __shared__ int s_wave[SCAN_WG / 64];
    const int base = blockIdx.x * SCAN_CHUNK;
    int sum = 0;
    for (int k = threadIdx.x; k < SCAN_CHUNK; k += SCAN_WG) {
        const int i = base + k;
        if (i < n) sum += in[i];
    }
    int total;
    block_scan_inclusive(sum, s_wave, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the chunk sums in place, total to *out_total.
__global__ __launch_bounds__(SCAN_WG) void k_scan_top(int* partial, int nparts, int* out_total, int* copy_total,
                                                      const int* flag) {
    if (!flag[0]) return;
    __shared__ int s_wave[SCAN_WG / 64];
    int carry = 0;
    for (int base = 0; base < nparts; base += SCAN_WG) {
        const int i = base + threadIdx.x;
        const int v = i < nparts ? partial[i] : 0;
        int total;
        const int inc = block_scan_inclusive(v, s_wave, &total);
        if (i < nparts) partial[i] = carry + inc - v;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out_total = carry;
        if (copy_total) *copy_total = carry;
    }
}

__global__ __launch_bounds__(SCAN_WG) void k_scan_down(int* in, int* out, int* copy, int n, int zero_input,
                                                       const int* __restrict__ partial, const int* flag) {
    if (!flag[0]) return;
    __shared__ int s_wave[SCAN_WG / 64];
    const int lo = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int sum = 0;
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = lo + k < n ? in[lo + k] : 0;
        sum += v[k];
    }
    int total;
    int run = partial[blockIdx.x] + block_scan_inclusive(sum, s_wave, &total) - sum;
    for (int k = 0; k < SCAN_PER; ++k) {
        const int i = lo + k;
        if (i < n) {
            out[i] = run;
            if (copy) copy[i] = run;
            if (zero_input) in[i] = 0;
        }
        run += v[k];
    }
}

// The same scan in one workgroup (chunks of SCAN_CHUNK with a running carry),
// for inputs of at most SCAN_SINGLE_MAX elements: one launch instead of three.
constexpr int SCAN_SINGLE_MAX = 8 * SCAN_CHUNK;

__device__ void scan_single_body(int* in, int* out, int* copy, int n, int zero_input) {
    __shared__ int s_wave[SCAN_WG / 64];
    int carry = 0;
    for (int base = 0; base < n; base += SCAN_CHUNK) {
        const int lo = base + threadIdx.x * SCAN_PER;
        int v[SCAN_PER];
        int sum = 0;
        for (int k = 0; k < SCAN_PER; ++k) {
            v[k] = lo + k < n ? in[lo + k] : 0;
            sum += v[k];
        }
        int total;
        int run = carry + block_scan_inclusive(sum, s_wave, &total) - sum;
        for (int k = 0; k < SCAN_PER; ++k) {
            const int i = lo + k;
            if (i < n) {
                out[i] = run;
                if (copy) copy[i] = run;
                if (zero_input) in[i] = 0;
            }
            run += v[k];
        }
        carry += total;
        __syncthreads();  // s_wave is reused by the next chunk
    }
    if (threadIdx.x == 0) {
        out[n] = carry;
        if (copy) copy[n] = carry;
    }
}

__global__ __launch_bounds__(SCAN_WG) void k_scan_single(int* in, int* out, int* copy, int n, int zero_input,
                                                         const int* flag) {
    if (!flag[0]) return;
    scan_single_body(in, out, copy, n, zero_input);
}

// k_unit_keys and the one-workgroup scan of the cell counts in one launch:
// the last workgroup to finish its keys (a completion counter, fenced on both
// sides so the counts of every workgroup are visible) scans the counts and
// resets the counter.  For at most SCAN_SINGLE_MAX cells.
__global__ __launch_bounds__(SCAN_WG) void k_unit_keys_scan(const double* __restrict__ pos,
                                                            const int* __restrict__ unit_ptr,
                                                            const int* __restrict__ unit_atoms,
                                                            const int* __restrict__ unit_sys, int nunits,
                                                            const SysInfo* __restrict__ sys,
                                                            const int* __restrict__ cell_rank, int* unit_key,
                                                            int* cell_count, int* cell_start, int* cell_cursor,
                                                            int ncells, int* done, const int* __restrict__ flag) {
    if (!flag[0]) return;
    const int u = blockIdx.x * SCAN_WG + threadIdx.x;
    if (u < nunits) unit_key_of(u, pos, unit_ptr, unit_atoms, unit_sys, sys, cell_rank, unit_key, cell_count, flag);
    __threadfence();
    __syncthreads();
    __shared__ int s_last;
    if (threadIdx.x == 0) s_last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (threadIdx.x == 0) *done = 0;
    scan_single_body(cell_count, cell_start, cell_cursor, ncells, 1);
}

__global__ void k_scatter_units(const int* __restrict__ unit_key, int nunits, int* cursor, int* sorted_units,
                                const int* flag) {
    if (!flag[0]) return;
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int slot = atomicAdd(&cursor[unit_key[u]], 1);
    sorted_units[slot] = u;
}

// Insertion sort of one cell's units by index (the scatter's atomics leave
// them in arbitrary order; the index order makes the layout deterministic).
__device__ __forceinline__ void sort_cell(int* sorted_units, int lo, int hi) {
    for (int i = lo + 1; i < hi; ++i) {
        int v = sorted_units[i];
        int j = i - 1;
        while (j >= lo && sorted_units[j] > v) {
            sorted_units[j + 1] = sorted_units[j];
            --j;
        }
        sorted_units[j + 1] = v;
    }
}

__global__ void k_sort_cells(const int* __restrict__ cell_start, int ncells, int* sorted_units,
                             const int* __restrict__ unit_ptr, int* unit_size_sorted, const int* flag) {
    if (!flag[0]) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int lo = cell_start[c], hi = cell_start[c + 1];
    sort_cell(sorted_units, lo, hi);
    if (unit_size_sorted)
        for (int i = lo; i < hi; ++i) {
            int u = sorted_units[i];
            unit_size_sorted[i] = unit_ptr[u + 1] - unit_ptr[u];
        }
}

// Places the atoms of unit u, at sorted position p.
__device__ __forceinline__ void place_unit(int p, int u, const int* __restrict__ unit_offset, int uniform,
                                           const int* __restrict__ unit_ptr, const int* __restrict__ unit_atoms,
                                           const int* __restrict__ unit_pos, const int* __restrict__ unit_sys,
                                           const SysInfo* __restrict__ sys, const int* __restrict__ excl_ptr,
                                           const int* __restrict__ excl_idx, int closed, int* sorted_atom,
                                           int* atom_slot, int2* excl_range, const int* flag) {
    const int si = unit_sys[u];
    if (!flag[1 + si]) return;  // same keys, same slots: nothing changes
    const SysInfo& S = sys[si];
    int off = (uniform > 0 ? p * uniform : unit_offset[p]) - S.atom_off + S.slot_off;
    for (int k = unit_ptr[u]; k < unit_ptr[u + 1]; ++k) {
        int a = unit_atoms[k];
        int slot = off + (k - unit_ptr[u]);
        sorted_atom[slot] = a;
        atom_slot[a] = slot;
        if (closed) {
            int lo = INT_MAX, hi = -1;
            for (int e = excl_ptr[a]; e < excl_ptr[a + 1]; ++e) {
                const int x = off + unit_pos[excl_idx[e]];
                lo = min(lo, x);
                hi = max(hi, x);
            }
            excl_range[slot] = make_int2(lo, hi);
        }
    }
}

// Places the atoms of sorted unit p at their slots.  Units are sorted by
// (system, cell), so the atoms before unit p are those of the earlier systems
// plus the earlier units of its own system; uniform > 0: every unit has that
// many atoms (no size scan).  When every exception partner of an atom lies in
// its own unit (closed units, e.g. rigid or flexible water), the slot range of
// its partners is written here as well.
__global__ void k_place_atoms(const int* __restrict__ sorted_units, const int* __restrict__ unit_offset, int uniform,
                              const int* __restrict__ unit_ptr, const int* __restrict__ unit_atoms,
                              const int* __restrict__ unit_pos, const int* __restrict__ unit_sys,
                              const SysInfo* __restrict__ sys, const int* __restrict__ excl_ptr,
                              const int* __restrict__ excl_idx, int closed, int nunits, int* sorted_atom,
                              int* atom_slot, int2* excl_range, const int* flag) {
    if (!flag[0]) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nunits) return;
    place_unit(p, sorted_units[p], unit_offset, uniform, unit_ptr, unit_atoms, unit_pos, unit_sys, sys, excl_ptr,
               excl_idx, closed, sorted_atom, atom_slot, excl_range, flag);
}

// Uniform units: one thread per unit of the scatter's order finds the unit's
// rank by index among the units of its cell (the position the index sort of
// the cell would give it; no size scan is needed) and places it, one launch
// instead of a sort and a placement.  The scatter's order is left as is.
__global__ void k_rank_place(const int* __restrict__ unit_key, const int* __restrict__ cell_start, int nunits,
                             const int* __restrict__ scattered, int uniform, const int* __restrict__ unit_ptr,
                             const int* __restrict__ unit_atoms, const int* __restrict__ unit_pos,
                             const int* __restrict__ unit_sys, const SysInfo* __restrict__ sys,
                             const int* __restrict__ excl_ptr, const int* __restrict__ excl_idx, int closed,
                             int* sorted_atom, int* atom_slot, int2* excl_range, const int* flag) {
    if (!flag[0]) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nunits) return;
    const int u = scattered[q];
    const int c = unit_key[u];
    const int lo = cell_start[c], hi = cell_start[c + 1];
    int p = lo;
    for (int k = lo; k < hi; ++k) p += scattered[k] < u ? 1 : 0;
    place_unit(p, u, nullptr, uniform, unit_ptr, unit_atoms, unit_pos, unit_sys, sys, excl_ptr, excl_idx, closed,
               sorted_atom, atom_slot, excl_range, flag);
}

// Slot range of each slot's exclusion partners once every atom is placed
// (units that do not hold all of their atoms' partners).
__global__ void k_excl_range(const int* __restrict__ sorted_atom, const int* __restrict__ atom_slot,
                             const int* __restrict__ excl_ptr, const int* __restrict__ excl_idx,
                             const int* __restrict__ block_sys, int npad, int2* excl_range, const int* flag) {
    if (!flag[0]) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    if (!flag[1 + block_sys[s / IBLOCK]]) return;
    int lo = INT_MAX, hi = -1;
    const int a = sorted_atom[s];
    if (a >= 0)
        for (int k = excl_ptr[a]; k < excl_ptr[a + 1]; ++k) {
            const int x = atom_slot[excl_idx[k]];
            lo = min(lo, x);
            hi = max(hi, x);
        }
    excl_range[s] = make_int2(lo, hi);
}

// Bounding box of superblock sb (SUPER consecutive blocks of system si) from
// its blocks' boxes; one wave, lane k reading block k.
__device__ void super_bounds_of(int sb, int si, const SysInfo& S, const int* __restrict__ super_off,
                                const double* center, const float* half, double* super_center, float* super_half) {
    const int lane = threadIdx.x & 63;
    const int first = S.block_off + (sb - super_off[si]) * SUPER;
    const int J = first + lane;
    double mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = 1e300;
        mx[k] = -1e300;
    }
    if (J < S.block_off + S.nblocks) {
        double d[3] = {center[3 * J] - center[3 * first], center[3 * J + 1] - center[3 * first + 1],
                       center[3 * J + 2] - center[3 * first + 2]};
        wrap_fast(d, S.box);
        for (int k = 0; k < 3; ++k) {
            mn[k] = d[k] - (double)half[3 * J + k];
            mx[k] = d[k] + (double)half[3 * J + k];
        }
    }
    for (int off = 1; off < 64; off <<= 1)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
        }
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            super_center[3 * sb + k] = center[3 * first + k] + 0.5 * (mn[k] + mx[k]);
            super_half[3 * sb + k] = (float)(0.5 * (mx[k] - mn[k]) * (1.0 + 1e-6) + 1e-6);
        }
}

// Two blocks per wave64: lanes 0-31 block 2*blockIdx.x, lanes 32-63 the next.
// The first slot of a block always holds an atom (padding only ends a
// system's last block).
__global__ __launch_bounds__(64) void k_block_bounds(const double* __restrict__ pos,
                                                     const int* __restrict__ sorted_atom, int nblocks,
                                                     const int* __restrict__ block_sys,
                                                     const SysInfo* __restrict__ sys, double reach,
                                                     double* center, float* half, int* minimg,
                                                     double* pos_build, float4* slot_pos, int* ntasks,
                                                     const int* flag) {
    if (!flag[0]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ntasks[0] = 0;  // k_clist appends the tasks anew
    int lane = threadIdx.x;
    int blk = blockIdx.x * 2 + (lane >> 5);
    const int si = blk < nblocks ? block_sys[blk] : 0;
    bool valid = blk < nblocks && flag[1 + si];  // unflagged systems keep their bounds
    int a = valid ? sorted_atom[blk * IBLOCK + (lane & 31)] : -1;
    int ref = valid ? sorted_atom[blk * IBLOCK] : 0;
    const Box& b = sys[si].box;
    double mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = 1e300;
        mx[k] = -1e300;
    }
    double d[3] = {0.0, 0.0, 0.0};
    if (a >= 0) {
        for (int k = 0; k < 3; ++k) d[k] = pos[3 * a + k] - pos[3 * ref + k];
        wrap_fast(d, b);
        for (int k = 0; k < 3; ++k) mn[k] = mx[k] = d[k];
        for (int k = 0; k < 3; ++k) pos_build[3 * a + k] = pos[3 * a + k];
    }
    for (int off = 16; off >= 1; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
        }
    if (valid) {
        float4 q = make_float4(1.0e18f, 1.0e18f, 1.0e18f, __int_as_float(-1));
        if (a >= 0)
            q = make_float4((float)(d[0] - 0.5 * (mn[0] + mx[0])), (float)(d[1] - 0.5 * (mn[1] + mx[1])),
                            (float)(d[2] - 0.5 * (mn[2] + mx[2])), __int_as_float(a));
        slot_pos[blk * IBLOCK + (lane & 31)] = q;
    }
    if ((lane & 31) == 0 && valid) {
        int need = b.triclinic;
        for (int k = 0; k < 3; ++k) {
            double c = pos[3 * ref + k] + 0.5 * (mn[k] + mx[k]);
            double h = 0.5 * (mx[k] - mn[k]);
            center[3 * blk + k] = c;
            half[3 * blk + k] = (float)(h * (1.0 + 1e-6) + 1e-6);
            if (h + reach >= 0.5 * b.v[k * 4] - 1e-6) need = 1;
        }
        minimg[blk] = need;
    }
}

// Superblock bounds of flagged systems, one wave per superblock.
__global__ __launch_bounds__(64) void k_super_bounds(int nsuper, const int* __restrict__ super_off, int nsys,
                                                     const SysInfo* __restrict__ sys, const double* __restrict__ center,
                                                     const float* __restrict__ half, double* super_center,
                                                     float* super_half, const int* flag) {
    if (!flag[0]) return;
    const int sb = blockIdx.x;
    const int si = find_sys(super_off, nsys, sb);
    if (!flag[1 + si]) return;
    super_bounds_of(sb, si, sys[si], super_off, center, half, super_center, super_half);
}


// Cluster lists, one wave per cluster C of block I = C / CPB (W clusters per
// workgroup).  Each pair of blocks of a system is taken by one
// of the two: in a system of nb blocks scanned flat, block I takes the nb/2
// blocks after it, cyclically (J = I + d mod nb for 0 < d < nb/2, and
// d = nb/2 from the first half), so every cluster scans and tests about the
// same number of blocks, wherever it sits in the sorted order; with
// superblocks, the blocks J > I.  Candidate blocks pass when their bounding
// boxes come within reach of the cluster's box; the atoms of the passing
// blocks and the later atoms of I itself are then tested against the
// cluster's box (or its atoms).  The list keeps the (sorted slot, atom) of
// the cluster's own atoms and of every atom of those blocks within reach
// (the pair kernel takes an entry of the cluster's own slots only after the
// i-atom), and the cluster appends its pair-kernel tasks: one run of up
// to TPT tiles of JTILE entries per task, packed as cluster * 256 + first
// tile (their order in the table does not matter, as forces are fixed-point
// sums; one atomic per workgroup reserves the runs of its clusters).  The
// atoms of four passing blocks are loaded together, so their loads overlap.
// Unflagged systems keep their lists and only append their tasks.
// Exclusion data the list build reads to mark excluded pairs per entry.
struct ExclLookup {
    const int* atom_slot;
    const int* ptr;
    const int* idx;
    const int2* range;
    // Bit k set: the atom at slot first + k is an exclusion partner of atom
    // a (at slot `slot`); every partner is checked, however many there are.
    // r: range[slot], loaded by the caller ahead of the distance tests.
    __device__ __forceinline__ unsigned bits(int2 r, int a, int first) const {
        if (r.y < first || r.x >= first + CLUSTER) return 0u;
        unsigned m = 0u;
        for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
            const int d = atom_slot[idx[k]] - first;
            if (d >= 0 && d < CLUSTER) m |= 1u << d;
        }
        return m;
    }
};

// The cluster's atoms a_i as FP32 offsets from its box centre, for the
// per-atom reach test of an offset d from that centre (valid where the
// nearest image of d to the centre is the nearest to every cluster atom
// within reach):  min_i |d - a_i|^2 = |d|^2 + min_i (|a_i|^2 - 2 d.a_i),
// three FMAs per atom.  The terms are wave-uniform (scalar registers).
struct ClusterAtoms {
    float mx2[CLUSTER], my2[CLUSTER], mz2[CLUSTER], aa[CLUSTER];  // -2 a_i, |a_i|^2
    bool on;          // apply the per-atom test on the FP64 path
    float reach2;     // (reach + 1e-4)^2: a margin far above the FP32 rounding
};

__device__ __forceinline__ float nearest2(float dx, float dy, float dz, const ClusterAtoms& ca) {
    float m = 3.0e38f;
#pragma unroll
    for (int i = 0; i < CLUSTER; ++i) m = fminf(m, fmaf(dz, ca.mz2[i], fmaf(dy, ca.my2[i], fmaf(dx, ca.mx2[i], ca.aa[i]))));
    return m + (dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ void append_entry(int slot, int atom, const double p[3], int first, const double cq[3],
                                             const double hq[3], const Box& b, double reach2, int cap,
                                             int2* out, unsigned char* out_ex, const ExclLookup& ex, int& count,
                                             unsigned long long below, const ClusterAtoms& ca) {
    const int2 r = slot >= 0 ? ex.range[slot] : make_int2(INT_MAX, -1);
    bool ok = false;
    if (atom >= 0) {
        if (slot >= first && slot < first + CLUSTER) {
            ok = true;
        } else {
            double d[3] = {p[0] - cq[0], p[1] - cq[1], p[2] - cq[2]};
            ok = box_gap2(d, hq, b) < reach2;
            if (ok && ca.on) {
                // Within reach of the box: keep the entry only if it is
                // within reach of one of the cluster's atoms.
                wrap_fast(d, b);
                ok = nearest2((float)d[0], (float)d[1], (float)d[2], ca) < ca.reach2;
            }
        }
    }
    const unsigned long long m = __ballot(ok);
    if (ok) {
        const int w = count + __popcll(m & below);
        if (w < cap) {
            out[w] = make_int2(slot, atom);
            out_ex[w] = (unsigned char)ex.bits(r, atom, first);
        }
    }
    count += __popcll(m);
}

// The same for a slot of a later block J whose offset from the cluster's
// centre is dc (J's centre, nearest image) + q.xyz (the slot's offset from
// J's centre), where every atom of J within reach of the cluster has that
// offset as its nearest image (J and the cluster fit in half the box with
// the reach).  The per-atom test alone decides (it implies the box test).
// r: ex.range[slot] (loaded by the caller with q).
__device__ __forceinline__ void append_near(int slot, float4 q, int2 r, float dcx, float dcy, float dcz, int first,
                                            int cap, int2* out, unsigned char* out_ex, const ExclLookup& ex,
                                            int& count, unsigned long long below, const ClusterAtoms& ca) {
    const int atom = __float_as_int(q.w);
    const bool ok = atom >= 0 && nearest2(dcx + q.x, dcy + q.y, dcz + q.z, ca) < ca.reach2;
    const unsigned long long m = __ballot(ok);
    if (ok) {
        const int w = count + __popcll(m & below);
        if (w < cap) {
            out[w] = make_int2(slot, atom);
            out_ex[w] = (unsigned char)ex.bits(r, atom, first);
        }
    }
    count += __popcll(m);
}

// Position of the k-th (from 0) set bit of m; m has more than k set bits.
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int k) {
    int pos = 0;
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll(m & ((1ull << w) - 1ull));
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += w;
        }
    }
    return pos;
}

// Clusters (one per wave) per workgroup of k_clist: two for a single system
// (small workgroups spread over every compute unit), sixteen for batches,
// where the workgroups of unflagged systems only append their tasks and one
// task reservation per workgroup keeps the atomics on the counter few.
constexpr int CLIST_WAVES_SINGLE = 2, CLIST_WAVES_BATCH = 16;

template <int W>
__global__ __launch_bounds__(64 * W) void k_clist(
    const double* __restrict__ pos, const int* __restrict__ sorted_atom, const int* __restrict__ block_sys,
    const SysInfo* __restrict__ sys, const double* __restrict__ center, const float* __restrict__ half,
    const int* __restrict__ super_off, const double* __restrict__ super_center,
    const float* __restrict__ super_half, const int* __restrict__ atom_slot, const int* __restrict__ excl_ptr,
    const int* __restrict__ excl_idx, const int2* __restrict__ excl_range, const int* __restrict__ minimg,
    const float4* __restrict__ slot_pos, double reach, int cap, int nclusters, int2* clist, unsigned char* clist_ex,
    int* ccount, int* tasks, int* ntasks, unsigned* status, const int* flag) {
    __shared__ int s_runs[W];
    __shared__ int s_base;
    if (!flag[0]) return;
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int C = blockIdx.x * W + wave;
    const int I = C / CPB;
    const int si = C < nclusters ? block_sys[I] : 0;
    const int first = C * CLUSTER;  // the cluster's first slot
    int count = 0;
    if (C < nclusters && !flag[1 + si]) {
        count = ccount[C];  // unflagged systems keep their lists
    } else if (C < nclusters && sorted_atom[first] >= 0) {
        const SysInfo& S = sys[si];
        const Box& b = S.box;
        const int jend = S.block_off + S.nblocks;
        const double reach2 = reach * reach;
        const unsigned long long below = (1ull << lane) - 1ull;
        int2* out = clist + (size_t)C * cap;
        unsigned char* out_ex = clist_ex + (size_t)C * cap;
        const ExclLookup ex{atom_slot, excl_ptr, excl_idx, excl_range};

        // The rest of block I from the cluster on: lanes 0 .. 31 - 8 q.
        const int rest = (I + 1) * IBLOCK - first;
        const int own_slot = lane < rest ? first + lane : -1;
        const int own_atom = own_slot >= 0 ? sorted_atom[own_slot] : -1;
        double own_p[3] = {0.0, 0.0, 0.0};
        if (own_atom >= 0)
            for (int k = 0; k < 3; ++k) own_p[k] = pos[3 * own_atom + k];
        // Bounding box of the cluster (lanes 0-7), relative to its first atom.
        const double rp[3] = {__shfl(own_p[0], 0, 64), __shfl(own_p[1], 0, 64), __shfl(own_p[2], 0, 64)};
        double mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = 1e300;
            mx[k] = -1e300;
        }
        if (lane < CLUSTER && own_atom >= 0) {
            double d[3] = {own_p[0] - rp[0], own_p[1] - rp[1], own_p[2] - rp[2]};
            wrap_fast(d, b);
            for (int k = 0; k < 3; ++k) mn[k] = mx[k] = d[k];
        }
        for (int off = 1; off < CLUSTER; off <<= 1)
            for (int k = 0; k < 3; ++k) {
                mn[k] = fmin(mn[k], __shfl_xor(mn[k], off, 64));
                mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
            }
        double cq[3], hq[3];
        for (int k = 0; k < 3; ++k) {
            const double lo = __shfl(mn[k], 0, 64), hi = __shfl(mx[k], 0, 64);
            cq[k] = rp[k] + 0.5 * (lo + hi);
            hq[k] = 0.5 * (hi - lo) * (1.0 + 1e-6) + 1e-6;
        }
        ClusterAtoms ca;
        ca.on = !minimg[I];
        ca.reach2 = (float)((reach + 1e-4) * (reach + 1e-4));
        {
            // Padding atoms sit far away and never pass.
            float o[3] = {1.0e18f, 1.0e18f, 1.0e18f};
            if (lane < CLUSTER && own_atom >= 0) {
                double d[3] = {own_p[0] - cq[0], own_p[1] - cq[1], own_p[2] - cq[2]};
                wrap_fast(d, b);
                for (int k = 0; k < 3; ++k) o[k] = (float)d[k];
            }
            const float oo = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
            const float m0 = -2.0f * o[0], m1 = -2.0f * o[1], m2 = -2.0f * o[2];
#pragma unroll
            for (int i = 0; i < CLUSTER; ++i) {
                ca.mx2[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m0), i));
                ca.my2[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m1), i));
                ca.mz2[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m2), i));
                ca.aa[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(oo), i));
            }
        }
        append_entry(own_slot, own_atom, own_p, first, cq, hq, b, reach2, cap, out, out_ex, ex, count, below, ca);

        // Blocks [J0, min(J0 + 64, jlim)) whose boxes come within reach.
        auto scan_blocks = [&](const int J0, const int jlim) {
            const int J = J0 + lane;
            bool pass = false, near = false;
            float dcf[3] = {0.0f, 0.0f, 0.0f};
            if (J < jlim) {
                double dc[3], h[3];
                for (int k = 0; k < 3; ++k) {
                    dc[k] = center[3 * J + k] - cq[k];
                    h[k] = hq[k] + (double)half[3 * J + k];
                }
                wrap_fast(dc, b);
                pass = box_gap2(dc, h, b) < reach2;
                near = !b.triclinic;
                for (int k = 0; k < 3; ++k) {
                    dcf[k] = (float)dc[k];
                    if (h[k] + reach >= 0.5 * b.v[k * 4] - 1e-4) near = false;
                }
            }
            const unsigned long long m = __ballot(pass);
            const int npass = __popcll(m);
            if (m != 0 && __ballot(pass && near) == m) {
                // Every passing block takes the slot-offset test, NEAR_ROUND
                // per round with all their loads issued together: lanes 0-31
                // the first of each pair, lanes 32-63 the second.
                constexpr int NEAR_ROUND = 8;
                const float4 none = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
                for (int r = 0; r < npass; r += NEAR_ROUND) {
                    int src[NEAR_ROUND / 2], sl[NEAR_ROUND / 2];
                    float4 q[NEAR_ROUND / 2];
                    int2 rg[NEAR_ROUND / 2];
#pragma unroll
                    for (int u = 0; u < NEAR_ROUND / 2; ++u) {
                        const int kk = r + 2 * u + (lane >> 5);
                        src[u] = kk < npass ? nth_set_bit(m, kk) : 0;
                        sl[u] = kk < npass ? (J0 + src[u]) * IBLOCK + (lane & 31) : -1;
                    }
#pragma unroll
                    for (int u = 0; u < NEAR_ROUND / 2; ++u) {
                        q[u] = sl[u] >= 0 ? slot_pos[sl[u]] : none;
                        rg[u] = sl[u] >= 0 ? excl_range[sl[u]] : make_int2(INT_MAX, -1);
                    }
#pragma unroll
                    for (int u = 0; u < NEAR_ROUND / 2; ++u)
                        append_near(sl[u], q[u], rg[u], __shfl(dcf[0], src[u], 64), __shfl(dcf[1], src[u], 64),
                                    __shfl(dcf[2], src[u], 64), first, cap, out, out_ex, ex, count, below, ca);
                }
                return;
            }
            // Four passing blocks per round: lanes 0-31 take the first of
            // each pair, lanes 32-63 the second.
            for (int r = 0; r < npass; r += 4) {
                const int ka = r + (lane >> 5), kb = r + 2 + (lane >> 5);
                const int slot_a = ka < npass ? (J0 + nth_set_bit(m, ka)) * IBLOCK + (lane & 31) : -1;
                const int slot_b = kb < npass ? (J0 + nth_set_bit(m, kb)) * IBLOCK + (lane & 31) : -1;
                const int atom_a = slot_a >= 0 ? sorted_atom[slot_a] : -1;
                const int atom_b = slot_b >= 0 ? sorted_atom[slot_b] : -1;
                double pa[3] = {0.0, 0.0, 0.0}, pb[3] = {0.0, 0.0, 0.0};
                if (atom_a >= 0)
                    for (int k = 0; k < 3; ++k) pa[k] = pos[3 * atom_a + k];
                if (atom_b >= 0)
                    for (int k = 0; k < 3; ++k) pb[k] = pos[3 * atom_b + k];
                append_entry(slot_a, atom_a, pa, first, cq, hq, b, reach2, cap, out, out_ex, ex, count, below, ca);
                append_entry(slot_b, atom_b, pb, first, cq, hq, b, reach2, cap, out, out_ex, ex, count, below, ca);
            }
        };
        if (S.nblocks > SUPER_MIN) {
            // Superblocks from I's own on, then the blocks of those in reach.
            const int sb0 = super_off[si], nsb = super_off[si + 1] - sb0;
            for (int B0 = (I - S.block_off) / SUPER; B0 < nsb; B0 += 64) {
                const int B = B0 + lane;
                bool pass = false;
                if (B < nsb) {
                    double dc[3], h[3];
                    for (int k = 0; k < 3; ++k) {
                        dc[k] = super_center[3 * (sb0 + B) + k] - cq[k];
                        h[k] = hq[k] + (double)super_half[3 * (sb0 + B) + k];
                    }
                    pass = box_gap2(dc, h, b) < reach2;
                }
                unsigned long long m = __ballot(pass);
                while (m) {
                    const int Bk = B0 + __ffsll((long long)m) - 1;
                    m &= m - 1;
                    const int lo = S.block_off + Bk * SUPER;
                    scan_blocks(max(lo, I + 1), min(lo + SUPER, jend));
                }
            }
        } else {
            const int nb = S.nblocks, rel = I - S.block_off;
            const int dmax = (nb - 1) / 2 + ((nb % 2 == 0 && rel < nb / 2) ? 1 : 0);
            const int fwd_end = min(I + 1 + dmax, jend);
            for (int J0 = I + 1; J0 < fwd_end; J0 += 64) scan_blocks(J0, fwd_end);
            const int wrap_end = S.block_off + max(0, I + 1 + dmax - jend);
            for (int J0 = S.block_off; J0 < wrap_end; J0 += 64) scan_blocks(J0, wrap_end);
        }
        if (count > cap) {
            if (lane == 0) atomicOr(status, ST_LIST_OVERFLOW);
            count = cap;
        }
        if (lane == 0) ccount[C] = count;
    } else if (C < nclusters) {
        if (lane == 0) ccount[C] = 0;  // a cluster of padding only
    }
    const int runs = ((count + JTILE - 1) / JTILE + TPT - 1) / TPT;
    if (lane == 0) s_runs[wave] = runs;
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 0;
        for (int w = 0; w < W; ++w) all += s_runs[w];
        s_base = all > 0 ? atomicAdd(ntasks, all) : 0;
    }
    __syncthreads();
    int base = s_base;
    for (int w = 0; w < wave; ++w) base += s_runs[w];
    for (int k = lane; k < runs; k += 64) tasks[base + k] = (C << 8) | (k * TPT);
}

// Position of cell (x, y, z) along a 3-D Hilbert curve over a 2^bits cube
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004:
// axes-to-transpose, then the transposed bits interleaved).  Consecutive
// cells on the curve share a face, so runs of sorted atoms stay compact.
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t h = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

inline int blocks_for(long long n, int t) { return (int)((n + t - 1) / t); }

void scan_async(NlistDev& nl, int* in, int* out, int* copy, int n, int zero_input, const int* flag, hipStream_t s) {
    if (n <= SCAN_SINGLE_MAX) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(SCAN_WG), 0, s, in, out, copy, n, zero_input, flag);
        return;
    }
    const int parts = blocks_for(n, SCAN_CHUNK);
    hipLaunchKernelGGL(k_scan_reduce, dim3(parts), dim3(SCAN_WG), 0, s, in, n, nl.scan_part.p, flag);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_WG), 0, s, nl.scan_part.p, parts, out + n,
                       copy ? copy + n : (int*)nullptr, flag);
    hipLaunchKernelGGL(k_scan_down, dim3(parts), dim3(SCAN_WG), 0, s, in, out, copy, n, zero_input, nl.scan_part.p,
                       flag);
}

}  // namespace

void nlist_setup(NlistDev& nl, const std::vector<NlistSystem>& systems, std::vector<SysInfo>& sys, double cutoff,
                 double skin, hipStream_t s) {
    const int nsys = (int)systems.size();
    nl.nsys = nsys;
    ++nl.generation;
    nl.cutoff = cutoff;
    nl.skin = skin;
    std::vector<int> eptr{0}, eidx, uptr{0}, uatoms, upos, usys, rank_all, block_sys, sorted, slot_of, atom_sys;
    std::vector<int> block_off{0}, chunk_off{0}, super_off{0};
    nl.use_super = false;
    int natoms = 0, nblocks = 0, nchunks = 0, ncells = 0, max_npad = 0;
    nl.uniform = -1;
    nl.closed = 1;
    for (int si = 0; si < nsys; ++si) {
        const NlistSystem& sy = systems[si];
        const int n = sy.n;
        if (n <= 0) throw Error(OMM_ERR_INVALID, "every system of a batch needs at least one particle");
        const std::vector<int>& exc1 = *sy.exc1;
        const std::vector<int>& exc2 = *sy.exc2;
        // Exclusion adjacency (both directions), global indices.
        std::vector<std::vector<int>> adj(n);
        for (size_t e = 0; e < exc1.size(); ++e) {
            adj[exc1[e]].push_back(exc2[e]);
            adj[exc2[e]].push_back(exc1[e]);
        }
        for (int i = 0; i < n; ++i) {
            for (int j : adj[i]) eidx.push_back(natoms + j);
            eptr.push_back((int)eidx.size());
        }
        // Sort units: connected components of the exception graph with <= 4
        // atoms move together; larger components are split into single atoms.
        std::vector<int> parent(n);
        std::iota(parent.begin(), parent.end(), 0);
        auto root = [&](int x) {
            while (parent[x] != x) x = parent[x] = parent[parent[x]];
            return x;
        };
        for (size_t e = 0; e < exc1.size(); ++e) {
            int a = root(exc1[e]), b2 = root(exc2[e]);
            if (a != b2) parent[std::max(a, b2)] = std::min(a, b2);
        }
        std::vector<std::vector<int>> comp(n);
        for (int i = 0; i < n; ++i) comp[root(i)].push_back(i);
        std::vector<int> unit_of(n, 0);
        upos.resize(natoms + n, 0);
        auto add_unit = [&](const std::vector<int>& atoms) {
            const int size = (int)atoms.size();
            nl.uniform = nl.uniform < 0 ? size : (nl.uniform == size ? size : 0);
            for (int k = 0; k < size; ++k) {
                upos[natoms + atoms[k]] = k;
                unit_of[atoms[k]] = (int)uptr.size() - 1;
                uatoms.push_back(natoms + atoms[k]);
            }
            uptr.push_back((int)uatoms.size());
            usys.push_back(si);
        };
        for (int i = 0; i < n; ++i) {
            if (comp[i].empty()) continue;
            if (comp[i].size() <= 4) {
                add_unit(comp[i]);
            } else {
                for (int a : comp[i]) add_unit(std::vector<int>{a});
            }
        }
        for (size_t e = 0; e < exc1.size(); ++e)
            if (unit_of[exc1[e]] != unit_of[exc2[e]]) nl.closed = 0;
        // Cell grid along each reciprocal direction, Hilbert ranks.
        SysInfo& S = sys[si];
        const Box& box = sy.box;
        for (int a = 0; a < 3; ++a) {
            double norm = std::sqrt(box.recip[a * 3] * box.recip[a * 3] + box.recip[a * 3 + 1] * box.recip[a * 3 + 1] +
                                    box.recip[a * 3 + 2] * box.recip[a * 3 + 2]);
            // A power of two per axis (cell edge 0.3-0.6 nm), where the
            // Hilbert curve never jumps; the cells only order the sort.
            const int want = std::max(1, std::min(512, (int)std::floor(1.0 / norm / 0.3)));
            int c = 1;
            while (2 * c <= want) c *= 2;
            S.ncell[a] = c;
        }
        const int nc = S.ncell[0] * S.ncell[1] * S.ncell[2];
        int bits = 1;
        while ((1 << bits) < std::max(S.ncell[0], std::max(S.ncell[1], S.ncell[2]))) ++bits;
        std::vector<std::pair<uint64_t, int>> codes(nc);
        for (int x = 0; x < S.ncell[0]; ++x)
            for (int y = 0; y < S.ncell[1]; ++y)
                for (int z = 0; z < S.ncell[2]; ++z) {
                    int lin = (x * S.ncell[1] + y) * S.ncell[2] + z;
                    codes[lin] = {hilbert3(x, y, z, bits), lin};
                }
        std::sort(codes.begin(), codes.end());
        std::vector<int> rank(nc);
        for (int r = 0; r < nc; ++r) rank[codes[r].second] = r;
        rank_all.insert(rank_all.end(), rank.begin(), rank.end());
        S.cell_off = ncells;
        ncells += nc;
        // Slots, blocks and PME chunks; the identity layout until the first build.
        const int nb = (n + IBLOCK - 1) / IBLOCK;
        S.atom_off = natoms;
        S.n = n;
        S.block_off = nblocks;
        S.nblocks = nb;
        S.slot_off = nblocks * IBLOCK;
        S.chunk_off = nchunks;
        for (int k = 0; k < nb * IBLOCK; ++k) {
            sorted.push_back(k < n ? natoms + k : -1);
            if (k < n) slot_of.push_back(S.slot_off + k);
        }
        for (int k = 0; k < nb; ++k) block_sys.push_back(si);
        const int nsb = (nb + SUPER - 1) / SUPER;
        super_off.push_back(super_off.back() + nsb);
        if (nb > SUPER_MIN) nl.use_super = true;
        atom_sys.insert(atom_sys.end(), n, si);
        natoms += n;
        nblocks += nb;
        nchunks += (nb * IBLOCK + PME_CHUNK - 1) / PME_CHUNK;
        block_off.push_back(nblocks);
        chunk_off.push_back(nchunks);
        max_npad = std::max(max_npad, nb * IBLOCK);
    }
    if (nl.uniform < 0) nl.uniform = 0;
    nl.natoms = natoms;
    nl.nblocks = nblocks;
    nl.nclusters = nblocks * CPB;
    nl.nslots = nblocks * IBLOCK;
    nl.nchunks = nchunks;
    nl.nunits = (int)uptr.size() - 1;
    nl.ncells = ncells;
    // List capacity: the atoms within reach of a compact cluster at liquid
    // density (a cluster's list holds about half of them), and never more
    // than the largest system; calibration and overflows grow it.
    double reach = cutoff + skin;
    double density = 100.0;
    double hb = 0.5 * std::cbrt(CLUSTER / density) * 1.6;
    double vol = std::pow(2 * hb, 3) + 6 * std::pow(2 * hb, 2) * reach + 3 * PI * 2 * hb * reach * reach +
                 4.0 / 3.0 * PI * reach * reach * reach;
    long long cap = std::isfinite(vol) ? (long long)std::min(density * vol, 1e9) + IBLOCK : (long long)max_npad;
    cap = std::min<long long>(cap, (long long)max_npad);
    cap = std::max<long long>(cap, 256);
    nl.cap = (int)((cap + JTILE - 1) / JTILE * JTILE);

    nl.unit_ptr.upload(uptr, s);
    nl.unit_atoms.upload(uatoms, s);
    nl.unit_pos.upload(upos, s);
    nl.unit_sys.upload(usys, s);
    nl.unit_key.alloc(nl.nunits);
    nl.cell_rank.upload(rank_all, s);
    std::vector<int> zeros(ncells, 0);
    nl.cell_count.upload(zeros, s);
    nl.cell_start.alloc(ncells + 1);
    nl.cell_cursor.alloc(ncells + 1);
    nl.sorted_units.alloc(nl.nunits);
    nl.unit_size_sorted.alloc(nl.nunits);
    nl.unit_offset.alloc(nl.nunits + 1);
    nl.sorted_atom.upload(sorted, s);
    nl.atom_slot.upload(slot_of, s);
    nl.block_sys.upload(block_sys, s);
    nl.block_off.upload(block_off, s);
    nl.chunk_off.upload(chunk_off, s);
    nl.excl_ptr.upload(eptr, s);
    if (eidx.empty()) eidx.push_back(0);
    nl.excl_idx.upload(eidx, s);
    nl.excl_range.alloc((size_t)nl.nslots);
    nl.slot_pos.alloc((size_t)nl.nslots);
    nl.block_center.alloc(3 * (size_t)nblocks);
    nl.block_half.alloc(3 * (size_t)nblocks);
    nl.block_minimg.alloc(nblocks);
    nl.nsuper = super_off.back();
    nl.super_off.upload(super_off, s);
    nl.super_center.alloc(3 * (size_t)nl.nsuper);
    nl.super_half.alloc(3 * (size_t)nl.nsuper);
    nl.done.alloc(1);
    fill_async<int>(nl.done.p, 1, 0, s);
    nl.pos_build.alloc(3 * (size_t)natoms);
    nl.cap = std::min(nl.cap, 255 * JTILE);
    nl.clist.alloc((size_t)nl.nclusters * nl.cap);
    nl.clist_ex.alloc((size_t)nl.nclusters * nl.cap);
    nl.tasks.alloc((size_t)nl.nclusters * ((nl.cap / JTILE + TPT - 1) / TPT));
    nl.ntasks.alloc(1);
    nl.ccount.alloc(nl.nclusters);
    nl.fslot_fx.alloc(3 * (size_t)nl.nslots);
    nl.fslot_d.alloc(3 * (size_t)nl.nslots);
    fill_async<long long>(nl.fslot_fx.p, 3 * (size_t)nl.nslots, 0ll, s);
    fill_async<double>(nl.fslot_d.p, 3 * (size_t)nl.nslots, 0.0, s);
    nl.scan_part.alloc((size_t)std::max(1, blocks_for(std::max(ncells, nl.nunits), SCAN_CHUNK)));
    std::vector<int> one(nsys + 1, 1);
    nl.flag.upload(one, s);
    nl.one.upload(one, s);
    nl.atom_sys.upload(atom_sys, s);
}

void nlist_resize(NlistDev& nl, int cap) {
    if (cap / JTILE > 255) throw Error(OMM_ERR_INVALID, "neighbour list needs more than 255 tiles per cluster");
    nl.cap = cap;
    nl.clist.alloc((size_t)nl.nclusters * nl.cap);
    nl.clist_ex.alloc((size_t)nl.nclusters * nl.cap);
    nl.tasks.alloc((size_t)nl.nclusters * ((nl.cap / JTILE + TPT - 1) / TPT));
}

void nlist_check(NlistDev& nl, const double* pos, int force, hipStream_t s) {
    double lim = 0.5 * nl.skin;
    hipLaunchKernelGGL(k_check, dim3(blocks_for(std::max(nl.natoms, nl.nsys), 256)), dim3(256), 0, s, pos,
                       nl.pos_build.p, nl.atom_sys.p, nl.natoms, nl.nsys, lim * lim, force, nl.flag.p);
    OMM_HIP(hipGetLastError());
}

void nlist_build(NlistDev& nl, const double* pos, const SysInfo* sys_dev, unsigned* status, hipStream_t s,
                 bool forced) {
    // Forced builds (the first step of every captured graph) run
    // unconditionally; otherwise every kernel exits early unless the
    // displacement check asked for a rebuild.
    const int* flag = forced ? nl.one.p : nl.flag.p;
    // Each kernel of the chain is a graph node that costs a few microseconds
    // even when it exits at once, so steps without a rebuild pay for the
    // chain's length: the cell-count scan rides on the key kernel and, for a
    // single system, the sort of uniform units on their placement.  (Batches
    // rebuild some replica on most steps, and there one thread per unit
    // places faster than one thread per cell.)
    if (nl.ncells <= SCAN_SINGLE_MAX) {
        hipLaunchKernelGGL(k_unit_keys_scan, dim3(blocks_for(nl.nunits, SCAN_WG)), dim3(SCAN_WG), 0, s, pos,
                           nl.unit_ptr.p, nl.unit_atoms.p, nl.unit_sys.p, nl.nunits, sys_dev, nl.cell_rank.p,
                           nl.unit_key.p, nl.cell_count.p, nl.cell_start.p, nl.cell_cursor.p, nl.ncells, nl.done.p,
                           flag);
    } else {
        hipLaunchKernelGGL(k_unit_keys, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, pos, nl.unit_ptr.p,
                           nl.unit_atoms.p, nl.unit_sys.p, nl.nunits, sys_dev, nl.cell_rank.p, nl.unit_key.p,
                           nl.cell_count.p, flag);
        scan_async(nl, nl.cell_count.p, nl.cell_start.p, nl.cell_cursor.p, nl.ncells, 1, flag, s);
    }
    hipLaunchKernelGGL(k_scatter_units, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, nl.unit_key.p, nl.nunits,
                       nl.cell_cursor.p, nl.sorted_units.p, flag);
    if (nl.uniform > 0 && nl.nsys == 1) {
        hipLaunchKernelGGL(k_rank_place, dim3(blocks_for(nl.nunits, 64)), dim3(64), 0, s, nl.unit_key.p,
                           nl.cell_start.p, nl.nunits, nl.sorted_units.p, nl.uniform, nl.unit_ptr.p, nl.unit_atoms.p,
                           nl.unit_pos.p, nl.unit_sys.p, sys_dev, nl.excl_ptr.p, nl.excl_idx.p, nl.closed,
                           nl.sorted_atom.p, nl.atom_slot.p, nl.excl_range.p, flag);
    } else {
        hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(nl.ncells, 256)), dim3(256), 0, s, nl.cell_start.p,
                           nl.ncells, nl.sorted_units.p, nl.unit_ptr.p,
                           nl.uniform > 0 ? (int*)nullptr : nl.unit_size_sorted.p, flag);
        if (nl.uniform == 0)
            scan_async(nl, nl.unit_size_sorted.p, nl.unit_offset.p, nullptr, nl.nunits, 0, flag, s);
        hipLaunchKernelGGL(k_place_atoms, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, nl.sorted_units.p,
                           nl.unit_offset.p, nl.uniform, nl.unit_ptr.p, nl.unit_atoms.p, nl.unit_pos.p,
                           nl.unit_sys.p, sys_dev, nl.excl_ptr.p, nl.excl_idx.p, nl.closed, nl.nunits,
                           nl.sorted_atom.p, nl.atom_slot.p, nl.excl_range.p, flag);
    }
    if (!nl.closed)
        hipLaunchKernelGGL(k_excl_range, dim3(blocks_for(nl.nslots, 256)), dim3(256), 0, s, nl.sorted_atom.p,
                           nl.atom_slot.p, nl.excl_ptr.p, nl.excl_idx.p, nl.block_sys.p, nl.nslots,
                           nl.excl_range.p, flag);
    double reach = nl.cutoff + nl.skin;
    hipLaunchKernelGGL(k_block_bounds, dim3(blocks_for(nl.nblocks, 2)), dim3(64), 0, s, pos, nl.sorted_atom.p,
                       nl.nblocks, nl.block_sys.p, sys_dev, reach, nl.block_center.p, nl.block_half.p,
                       nl.block_minimg.p, nl.pos_build.p, nl.slot_pos.p, nl.ntasks.p, flag);
    if (nl.use_super)
        hipLaunchKernelGGL(k_super_bounds, dim3(nl.nsuper), dim3(64), 0, s, nl.nsuper, nl.super_off.p, nl.nsys,
                           sys_dev, nl.block_center.p, nl.block_half.p, nl.super_center.p, nl.super_half.p, flag);
    auto clist = [&](auto kernel, int w) {
        hipLaunchKernelGGL(kernel, dim3(blocks_for(nl.nclusters, w)), dim3(64 * w), 0, s, pos, nl.sorted_atom.p,
                           nl.block_sys.p, sys_dev, nl.block_center.p, nl.block_half.p, nl.super_off.p,
                           nl.super_center.p, nl.super_half.p, nl.atom_slot.p, nl.excl_ptr.p, nl.excl_idx.p,
                           nl.excl_range.p, nl.block_minimg.p, nl.slot_pos.p, reach, nl.cap, nl.nclusters, nl.clist.p,
                           nl.clist_ex.p, nl.ccount.p, nl.tasks.p, nl.ntasks.p, status, flag);
    };
    if (nl.nsys == 1)
        clist(k_clist<CLIST_WAVES_SINGLE>, CLIST_WAVES_SINGLE);
    else
        clist(k_clist<CLIST_WAVES_BATCH>, CLIST_WAVES_BATCH);
    OMM_HIP(hipGetLastError());
}

}  // namespace ommhip
