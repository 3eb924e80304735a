// Tile neighbour list of the direct-space pair kernel, for a batch of
// independent systems built by one set of launches.
//
// Atoms are grouped into sort units (a small exception-connected molecule, or
// a single atom), binned by the fractional coordinates of the unit's first
// atom on a per-system cell grid (power-of-two cells per axis, 0.3-0.6 nm),
// and counting-sorted by (system, Hilbert rank of the cell); ties inside a
// cell are broken by unit index so the order (and hence every FP32 partial
// sum downstream) is deterministic and the same as for the system built
// alone.  Consecutive runs
// of IBLOCK sorted atoms of one system form blocks with an FP64 centre and
// FP32 half-extents; each block holds four 8-atom clusters.  The list of a
// cluster holds, in increasing order, the (sorted slot, atom) of its own
// atoms and of the later atoms of the same system within cutoff + skin of
// its bounding box; blocks whose boxes are out of reach are skipped whole
// (64-lane ballot + popcount compaction).  The pair kernel walks the cluster
// lists in tiles of 8 i-atoms x 64 entries.  Pairs inside a list are kept
// with j_slot > i_slot, so each unordered pair is visited once.
#pragma once

#include "common.h"

namespace ommhip {

constexpr int IBLOCK = 32;  // atoms per block (the unit of the coarse list)
constexpr int CLUSTER = 8;  // i-atoms per cluster (the unit of the pair kernel's lists)
constexpr int CPB = IBLOCK / CLUSTER;  // clusters per block
constexpr int SUPER = 64;   // blocks per superblock (the coarse level of the candidate search)
constexpr int SUPER_MIN = 1024;  // systems with more blocks search superblocks first
constexpr int JTILE = 64;   // j-slots per tile
constexpr int TPT = 4;      // tiles of one cluster per pair-kernel task
constexpr int PME_CHUNK = 64;  // sorted slots per PME spread/gather workgroup

// Squared displacement of an atom from its position at the last list build,
// summed x, y, z with fused multiply-adds.  The integrators and the list's own
// displacement check both use this one form, so the same positions raise the
// same rebuild requests wherever the test runs.
__host__ __device__ __forceinline__ double skin_d2(double dx, double dy, double dz) {
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// Host description of one system for the list set-up (local atom indices).
struct NlistSystem {
    int n = 0;
    const std::vector<int>* exc1 = nullptr;
    const std::vector<int>* exc2 = nullptr;
    Box box{};
};

struct NlistDev {
    int nsys = 0;
    int generation = 0;  // advanced by every nlist_setup (slot layout changes)
    int natoms = 0, nslots = 0, nblocks = 0, nclusters = 0, nchunks = 0, cap = 0, nunits = 0, ncells = 0;
    int uniform = 0;  // atoms per unit when all units are the same size, else 0
    int closed = 0;   // every exception pair lies inside one unit
    double cutoff = 0, skin = 0;
    DevBuf<int> unit_ptr, unit_atoms, unit_pos, unit_sys, unit_key, cell_rank, cell_count, cell_start, cell_cursor;
    DevBuf<int> sorted_units, unit_size_sorted, unit_offset;
    DevBuf<int> sorted_atom, atom_slot, block_sys;
    DevBuf<int> block_off, chunk_off;  // per-system prefix tables (nsys + 1)
    // Exclusion partners (exception pairs, both directions) by atom, and per
    // sorted slot the lowest and highest slot of its atom's partners ({INT_MAX,
    // -1} for none), which lets the list build skip entries that cannot hold
    // an exclusion with a cluster.
    DevBuf<int> excl_ptr, excl_idx;
    DevBuf<int2> excl_range;
    DevBuf<double> block_center, pos_build;
    DevBuf<float> block_half;
    // Per sorted slot: the FP32 offset of its atom from its block's centre
    // (minimum image within the block), .w the atom (int bits); padding
    // slots sit far away with atom -1.  Written with the block bounds.
    DevBuf<float4> slot_pos;
    DevBuf<int> block_minimg;
    // Superblocks of SUPER consecutive blocks (systems with more than
    // SUPER_MIN blocks): bounding boxes and per-system offsets.
    int nsuper = 0;
    bool use_super = false;
    DevBuf<int> super_off;
    DevBuf<double> super_center;
    DevBuf<float> super_half;
    // Completion counter of the merged key-and-scan kernel (zero between builds).
    DevBuf<int> done;
    DevBuf<int2> clist;         // cluster lists [nclusters][cap] of (sorted slot, atom)
    // Per list entry, bit k: the cluster's k-th atom and the entry's atom are
    // an excluded pair (any number of exclusions per atom).
    DevBuf<unsigned char> clist_ex;
    DevBuf<int> ccount;
    DevBuf<int> tasks, ntasks;  // pair-kernel tasks (cluster << 8 | first tile) and their count
    // Direct-space forces by sorted slot, one component after another
    // ([3][nslots]): the pair kernel's adds from consecutive list entries fall
    // on consecutive addresses.  Zero between evaluations.
    DevBuf<long long> fslot_fx;
    DevBuf<double> fslot_d;
    DevBuf<int> scan_part;  // chunk sums of the multi-workgroup scans
    DevBuf<int> atom_sys;
    // Rebuild requests: [0] some system, [1 + s] system s.  Systems that are
    // not flagged keep their order, bounds and lists through a build, so a
    // replica rebuilds exactly when it would alone.
    DevBuf<int> flag;
    DevBuf<int> one;  // all ones: the flags of forced builds
};

// Host-side set-up: units from each system's exception graph, Hilbert cell
// ranks, slot/block/cell ranges (written into sys[s]), the identity slot
// layout (so PME may run before the first build), sizes.
void nlist_setup(NlistDev& nl, const std::vector<NlistSystem>& systems, std::vector<SysInfo>& sys, double cutoff,
                 double skin, hipStream_t s);

// New per-cluster list capacity (reallocates the lists and the task table).
void nlist_resize(NlistDev& nl, int cap);

// Requests a rebuild when an atom moved more than skin/2 since the last build
// (or always when force != 0).  Device-side: usable inside a captured graph.
void nlist_check(NlistDev& nl, const double* pos, int force, hipStream_t s);

// Rebuilds the list if requested (every kernel exits early otherwise), or
// unconditionally when forced.  The pair kernel, the next kernel to run on
// the list, clears the requests.
void nlist_build(NlistDev& nl, const double* pos, const SysInfo* sys_dev, unsigned* status, hipStream_t s,
                 bool forced = false);

}  // namespace ommhip
