// Tile neighbour list: deterministic spatial sort, block bounds, ballot-compacted
// j-lists and the task prefix sum of the pair kernel.
#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>

#include "nlist.h"

namespace ommhip {

namespace {

constexpr int SCAN_THREADS = 1024;

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

__global__ void k_check(const double* __restrict__ pos, const double* __restrict__ pos_build, int n,
                        double lim2, int force, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (i < n) {
        double dx = pos[3 * i] - pos_build[3 * i];
        double dy = pos[3 * i + 1] - pos_build[3 * i + 1];
        double dz = pos[3 * i + 2] - pos_build[3 * i + 2];
        double d2 = dx * dx + dy * dy + dz * dz;
        moved = !(d2 <= lim2);  // NaN also requests a rebuild
    }
    if (force && i == 0) moved = true;
    if (__ballot(moved) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void k_zero_i32(int* p, int n, const int* flag) {
    if (!flag[0]) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_unit_keys(const double* __restrict__ pos, const int* __restrict__ unit_ptr,
                            const int* __restrict__ unit_atoms, int nunits, const Box* __restrict__ box,
                            int cx, int cy, int cz, const int* __restrict__ cell_rank, int* unit_key,
                            int* cell_count, const int* flag) {
    if (!flag[0]) return;
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int a = unit_atoms[unit_ptr[u]];
    const Box& b = *box;
    int c[3];
    int nc[3] = {cx, cy, cz};
    for (int ax = 0; ax < 3; ++ax) {
        double f = pos[3 * a] * b.recip[ax * 3] + pos[3 * a + 1] * b.recip[ax * 3 + 1] +
                   pos[3 * a + 2] * b.recip[ax * 3 + 2];
        f -= floor(f);
        int v = (int)(f * nc[ax]);
        c[ax] = v < 0 ? 0 : (v >= nc[ax] ? nc[ax] - 1 : v);
    }
    int key = cell_rank[(c[0] * cy + c[1]) * cz + c[2]];
    unit_key[u] = key;
    atomicAdd(&cell_count[key], 1);
}

// Exclusive scan of in[0..n) into out[0..n], out[n] = total (one workgroup).
// copy (optional) receives the same values.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const int* __restrict__ in, int* out, int* copy, int n,
                                                       const int* flag) {
    if (!flag[0]) return;
    __shared__ int s_part[SCAN_THREADS];
    int t = threadIdx.x;
    int per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
    int lo = min(n, t * per), hi = min(n, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += in[i];
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        int v = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = s_part[t] - sum;
    for (int i = lo; i < hi; ++i) {
        out[i] = run;
        if (copy) copy[i] = run;
        run += in[i];
    }
    if (t == SCAN_THREADS - 1) {
        out[n] = s_part[t];
        if (copy) copy[n] = s_part[t];
    }
}

__global__ void k_scatter_units(const int* __restrict__ unit_key, int nunits, int* cursor, int* sorted_units,
                                const int* flag) {
    if (!flag[0]) return;
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    int slot = atomicAdd(&cursor[unit_key[u]], 1);
    sorted_units[slot] = u;
}

__global__ void k_sort_cells(const int* __restrict__ cell_start, int ncells, int* sorted_units,
                             const int* __restrict__ unit_ptr, int* unit_size_sorted, const int* flag) {
    if (!flag[0]) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int lo = cell_start[c], hi = cell_start[c + 1];
    for (int i = lo + 1; i < hi; ++i) {
        int v = sorted_units[i];
        int j = i - 1;
        while (j >= lo && sorted_units[j] > v) {
            sorted_units[j + 1] = sorted_units[j];
            --j;
        }
        sorted_units[j + 1] = v;
    }
    for (int i = lo; i < hi; ++i) {
        int u = sorted_units[i];
        unit_size_sorted[i] = unit_ptr[u + 1] - unit_ptr[u];
    }
}

__global__ void k_place_atoms(const int* __restrict__ sorted_units, const int* __restrict__ unit_offset,
                              const int* __restrict__ unit_ptr, const int* __restrict__ unit_atoms, int nunits,
                              int* sorted_atom, int* atom_slot, const int* flag) {
    if (!flag[0]) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nunits) return;
    int u = sorted_units[s];
    int off = unit_offset[s];
    for (int k = unit_ptr[u]; k < unit_ptr[u + 1]; ++k) {
        int a = unit_atoms[k];
        int slot = off + (k - unit_ptr[u]);
        sorted_atom[slot] = a;
        atom_slot[a] = slot;
    }
}

__global__ void k_excl_sorted(const int* __restrict__ sorted_atom, const int* __restrict__ atom_slot,
                              const int* __restrict__ excl_ptr, const int* __restrict__ excl_idx, int npad,
                              int* excl_sorted, const int* flag) {
    if (!flag[0]) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    int v[MAXEX];
    for (int k = 0; k < MAXEX; ++k) v[k] = INT_MAX;
    int a = sorted_atom[s];
    if (a >= 0) {
        int cnt = 0;
        for (int k = excl_ptr[a]; k < excl_ptr[a + 1] && cnt < MAXEX; ++k) {
            int x = atom_slot[excl_idx[k]];
            int j = cnt++;
            while (j > 0 && v[j - 1] > x) {
                v[j] = v[j - 1];
                --j;
            }
            v[j] = x;
        }
    }
    for (int k = 0; k < MAXEX; ++k) excl_sorted[(size_t)s * MAXEX + k] = v[k];
}

// Two blocks per wave64: lanes 0-31 block 2*blockIdx.x, lanes 32-63 the next.
__global__ __launch_bounds__(64) void k_block_bounds(const double* __restrict__ pos,
                                                     const int* __restrict__ sorted_atom, int n, int nblocks,
                                                     const Box* __restrict__ box, double reach,
                                                     double* center, float* half, int* minimg,
                                                     double* pos_build, const int* flag) {
    if (!flag[0]) return;
    const Box& b = *box;
    int lane = threadIdx.x;
    int blk = blockIdx.x * 2 + (lane >> 5);
    int slot = blk * IBLOCK + (lane & 31);
    bool live = blk < nblocks && slot < n;
    int a = live ? sorted_atom[slot] : -1;
    int ref = blk < nblocks ? sorted_atom[blk * IBLOCK] : 0;
    double mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = 1e300;
        mx[k] = -1e300;
    }
    if (a >= 0) {
        double d[3] = {pos[3 * a] - pos[3 * ref], pos[3 * a + 1] - pos[3 * ref + 1], pos[3 * a + 2] - pos[3 * ref + 2]};
        wrap_fast(d, b);
        for (int k = 0; k < 3; ++k) mn[k] = mx[k] = d[k];
        for (int k = 0; k < 3; ++k) pos_build[3 * a + k] = pos[3 * a + k];
    }
    for (int off = 16; off >= 1; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
        }
    if ((lane & 31) == 0 && blk < nblocks) {
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

__global__ __launch_bounds__(64) void k_jlist(const double* __restrict__ pos, const int* __restrict__ sorted_atom,
                                              int n, int nblocks, const double* __restrict__ center,
                                              const float* __restrict__ half, const Box* __restrict__ box,
                                              double reach, int cap, int* jlist, int* jcount, unsigned* status,
                                              const int* flag) {
    if (!flag[0]) return;
    const Box& b = *box;
    const int I = blockIdx.x;
    const int lane = threadIdx.x;
    const double reach2 = reach * reach;
    double cI[3] = {center[3 * I], center[3 * I + 1], center[3 * I + 2]};
    double hI[3] = {half[3 * I], half[3 * I + 1], half[3 * I + 2]};
    int* out = jlist + (size_t)I * cap;
    int own = min(IBLOCK, n - I * IBLOCK);
    if (lane < own) out[lane] = I * IBLOCK + lane;
    int count = own;
    for (int J0 = I + 1; J0 < nblocks; J0 += 64) {
        int J = J0 + lane;
        bool pass = false;
        if (J < nblocks) {
            double dc[3], h[3];
            for (int k = 0; k < 3; ++k) {
                dc[k] = center[3 * J + k] - cI[k];
                h[k] = hI[k] + (double)half[3 * J + k];
            }
            pass = box_gap2(dc, h, b) < reach2;
        }
        unsigned long long m = __ballot(pass);
        while (m) {
            int j1 = __ffsll((long long)m) - 1;
            m &= m - 1;
            int j2 = -1;
            if (m) {
                j2 = __ffsll((long long)m) - 1;
                m &= m - 1;
            }
            int Jx = (lane < 32) ? J0 + j1 : (j2 >= 0 ? J0 + j2 : -1);
            int slot = Jx >= 0 ? Jx * IBLOCK + (lane & 31) : -1;
            bool ok = false;
            if (slot >= 0 && slot < n) {
                int a = sorted_atom[slot];
                double d[3] = {pos[3 * a] - cI[0], pos[3 * a + 1] - cI[1], pos[3 * a + 2] - cI[2]};
                ok = box_gap2(d, hI, b) < reach2;
            }
            unsigned long long mm = __ballot(ok);
            int off = __popcll(mm & ((1ull << lane) - 1ull));
            if (ok) {
                int w = count + off;
                if (w < cap) out[w] = slot;
            }
            count += __popcll(mm);
        }
    }
    if (lane == 0) {
        if (count > cap) {
            atomicOr(status, ST_LIST_OVERFLOW);
            count = cap;
        }
        jcount[I] = count;
    }
}

__global__ void k_tiles(const int* __restrict__ jcount, int nblocks, int* tiles, const int* flag) {
    if (!flag[0]) return;
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I < nblocks) tiles[I] = (jcount[I] + JTILE - 1) / JTILE;
}

__global__ void k_clear_flag(int* flag) { flag[0] = 0; }

uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    auto spread = [](uint32_t v) {
        uint64_t r = 0;
        for (int i = 0; i < 10; ++i) r |= (uint64_t)((v >> i) & 1u) << (3 * i);
        return (uint32_t)r;
    };
    return (spread(x) << 2) | (spread(y) << 1) | spread(z);
}

inline int blocks_for(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

void nlist_setup(NlistDev& nl, int n, const std::vector<int>& exc1, const std::vector<int>& exc2, double cutoff,
                 double skin, const Box& box, hipStream_t s) {
    nl.n = n;
    nl.nblocks = (n + IBLOCK - 1) / IBLOCK;
    nl.npad = nl.nblocks * IBLOCK;
    nl.cutoff = cutoff;
    nl.skin = skin;
    // Exclusion adjacency (both directions) in original indices.
    std::vector<std::vector<int>> adj(n);
    for (size_t e = 0; e < exc1.size(); ++e) {
        adj[exc1[e]].push_back(exc2[e]);
        adj[exc2[e]].push_back(exc1[e]);
    }
    std::vector<int> eptr(n + 1, 0), eidx;
    for (int i = 0; i < n; ++i) {
        if ((int)adj[i].size() > MAXEX)
            throw Error(OMM_ERR_INVALID, "device path supports at most " + std::to_string(MAXEX) +
                                             " exceptions per particle");
        eptr[i + 1] = eptr[i] + (int)adj[i].size();
        eidx.insert(eidx.end(), adj[i].begin(), adj[i].end());
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
    std::vector<int> uptr{0}, uatoms;
    for (int i = 0; i < n; ++i) {
        if (comp[i].empty()) continue;
        if (comp[i].size() <= 4) {
            uatoms.insert(uatoms.end(), comp[i].begin(), comp[i].end());
            uptr.push_back((int)uatoms.size());
        } else {
            for (int a : comp[i]) {
                uatoms.push_back(a);
                uptr.push_back((int)uatoms.size());
            }
        }
    }
    nl.nunits = (int)uptr.size() - 1;
    // Cell grid of ~0.35 nm along each reciprocal direction.
    for (int a = 0; a < 3; ++a) {
        double norm = std::sqrt(box.recip[a * 3] * box.recip[a * 3] + box.recip[a * 3 + 1] * box.recip[a * 3 + 1] +
                                box.recip[a * 3 + 2] * box.recip[a * 3 + 2]);
        double width = 1.0 / norm;
        nl.ncell[a] = std::max(1, std::min(1023, (int)std::floor(width / 0.35)));
    }
    nl.ncells = nl.ncell[0] * nl.ncell[1] * nl.ncell[2];
    std::vector<std::pair<uint32_t, int>> codes(nl.ncells);
    for (int x = 0; x < nl.ncell[0]; ++x)
        for (int y = 0; y < nl.ncell[1]; ++y)
            for (int z = 0; z < nl.ncell[2]; ++z) {
                int lin = (x * nl.ncell[1] + y) * nl.ncell[2] + z;
                codes[lin] = {morton3(x, y, z), lin};
            }
    std::sort(codes.begin(), codes.end());
    std::vector<int> rank(nl.ncells);
    for (int r = 0; r < nl.ncells; ++r) rank[codes[r].second] = r;
    // List capacity: the atoms within reach of a compact block at liquid
    // density, doubled, and never less than all atoms when the box is small.
    double reach = cutoff + skin;
    double density = 100.0;
    double hb = 0.5 * std::cbrt(IBLOCK / density) * 1.6;
    double vol = std::pow(2 * hb, 3) + 6 * std::pow(2 * hb, 2) * reach + 3 * PI * 2 * hb * reach * reach +
                 4.0 / 3.0 * PI * reach * reach * reach;
    long long cap = (long long)(2.0 * density * vol) + IBLOCK;
    cap = std::min<long long>(cap, (long long)nl.npad);
    cap = std::max<long long>(cap, 256);
    nl.cap = (int)((cap + JTILE - 1) / JTILE * JTILE);

    nl.unit_ptr.upload(uptr, s);
    nl.unit_atoms.upload(uatoms, s);
    nl.unit_key.alloc(nl.nunits);
    nl.cell_rank.upload(rank, s);
    nl.cell_count.alloc(nl.ncells);
    nl.cell_start.alloc(nl.ncells + 1);
    nl.cell_cursor.alloc(nl.ncells + 1);
    nl.sorted_units.alloc(nl.nunits);
    nl.unit_size_sorted.alloc(nl.nunits);
    nl.unit_offset.alloc(nl.nunits + 1);
    std::vector<int> pad(nl.npad, -1);
    nl.sorted_atom.upload(pad, s);
    nl.atom_slot.alloc(n);
    nl.excl_ptr.upload(eptr, s);
    if (eidx.empty()) eidx.push_back(0);
    nl.excl_idx.upload(eidx, s);
    nl.excl_sorted.alloc((size_t)nl.npad * MAXEX);
    nl.block_center.alloc(3 * (size_t)nl.nblocks);
    nl.block_half.alloc(3 * (size_t)nl.nblocks);
    nl.block_minimg.alloc(nl.nblocks);
    nl.pos_build.alloc(3 * (size_t)n);
    nl.jlist.alloc((size_t)nl.nblocks * nl.cap);
    nl.jcount.alloc(nl.nblocks);
    nl.tiles.alloc(nl.nblocks);
    nl.task_start.alloc(nl.nblocks + 1);
    std::vector<int> one{1};
    nl.flag.upload(one, s);
}

void nlist_check(NlistDev& nl, const double* pos, const Box* box_dev, int force, hipStream_t s) {
    (void)box_dev;
    double lim = 0.5 * nl.skin;
    hipLaunchKernelGGL(k_check, dim3(blocks_for(nl.n, 256)), dim3(256), 0, s, pos, nl.pos_build.p, nl.n, lim * lim,
                       force, nl.flag.p);
    OMM_HIP(hipGetLastError());
}

void nlist_build(NlistDev& nl, const double* pos, const Box* box_dev, unsigned* status, hipStream_t s) {
    const int* flag = nl.flag.p;
    hipLaunchKernelGGL(k_zero_i32, dim3(std::min(1024, blocks_for(nl.ncells, 256))), dim3(256), 0, s, nl.cell_count.p,
                       nl.ncells, flag);
    hipLaunchKernelGGL(k_unit_keys, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, pos, nl.unit_ptr.p,
                       nl.unit_atoms.p, nl.nunits, box_dev, nl.ncell[0], nl.ncell[1], nl.ncell[2], nl.cell_rank.p,
                       nl.unit_key.p, nl.cell_count.p, flag);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, s, nl.cell_count.p, nl.cell_start.p, nl.cell_cursor.p,
                       nl.ncells, flag);
    hipLaunchKernelGGL(k_scatter_units, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, nl.unit_key.p, nl.nunits,
                       nl.cell_cursor.p, nl.sorted_units.p, flag);
    hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(nl.ncells, 256)), dim3(256), 0, s, nl.cell_start.p, nl.ncells,
                       nl.sorted_units.p, nl.unit_ptr.p, nl.unit_size_sorted.p, flag);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, s, nl.unit_size_sorted.p, nl.unit_offset.p,
                       (int*)nullptr, nl.nunits, flag);
    hipLaunchKernelGGL(k_place_atoms, dim3(blocks_for(nl.nunits, 256)), dim3(256), 0, s, nl.sorted_units.p,
                       nl.unit_offset.p, nl.unit_ptr.p, nl.unit_atoms.p, nl.nunits, nl.sorted_atom.p,
                       nl.atom_slot.p, flag);
    hipLaunchKernelGGL(k_excl_sorted, dim3(blocks_for(nl.npad, 256)), dim3(256), 0, s, nl.sorted_atom.p,
                       nl.atom_slot.p, nl.excl_ptr.p, nl.excl_idx.p, nl.npad, nl.excl_sorted.p, flag);
    double reach = nl.cutoff + nl.skin;
    hipLaunchKernelGGL(k_block_bounds, dim3(blocks_for(nl.nblocks, 2)), dim3(64), 0, s, pos, nl.sorted_atom.p, nl.n,
                       nl.nblocks, box_dev, reach, nl.block_center.p, nl.block_half.p, nl.block_minimg.p,
                       nl.pos_build.p, flag);
    hipLaunchKernelGGL(k_jlist, dim3(nl.nblocks), dim3(64), 0, s, pos, nl.sorted_atom.p, nl.n, nl.nblocks,
                       nl.block_center.p, nl.block_half.p, box_dev, reach, nl.cap, nl.jlist.p, nl.jcount.p, status,
                       flag);
    hipLaunchKernelGGL(k_tiles, dim3(blocks_for(nl.nblocks, 256)), dim3(256), 0, s, nl.jcount.p, nl.nblocks,
                       nl.tiles.p, flag);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, s, nl.tiles.p, nl.task_start.p, (int*)nullptr,
                       nl.nblocks, flag);
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, nl.flag.p);
    OMM_HIP(hipGetLastError());
}

}  // namespace ommhip
