// Tile neighbour list of the direct-space pair kernel.
//
// Atoms are grouped into sort units (a small exception-connected molecule, or
// a single atom), binned by the fractional coordinates of the unit's first
// atom on a cell grid of ~0.35 nm, and counting-sorted in Morton order of the
// cells; ties inside a cell are broken by unit index so the order (and hence
// every FP32 partial sum downstream) is deterministic.  Consecutive runs of
// IBLOCK sorted atoms form blocks with an FP64 centre and FP32 half-extents.
// For every block I the list holds, in increasing order, the sorted indices
// of the atoms of blocks J >= I within cutoff + skin of I's bounding box
// (64-lane ballot + mbcnt compaction).  Pairs inside a list are kept with
// j_sorted > i_sorted, so each unordered pair is visited once.
#pragma once

#include "common.h"

namespace ommhip {

constexpr int IBLOCK = 32;  // i-atoms per tile (two j half-tiles per wave64)
constexpr int JTILE = 64;   // j-slots per tile
constexpr int MAXEX = 8;    // exclusion partners per atom handled by the pair kernel

struct NlistDev {
    int n = 0, npad = 0, nblocks = 0, cap = 0, nunits = 0, ncells = 0;
    int ncell[3] = {1, 1, 1};
    double cutoff = 0, skin = 0;
    DevBuf<int> unit_ptr, unit_atoms, unit_key, cell_rank, cell_count, cell_start, cell_cursor;
    DevBuf<int> sorted_units, unit_size_sorted, unit_offset;
    DevBuf<int> sorted_atom, atom_slot;
    DevBuf<int> excl_ptr, excl_idx, excl_sorted;
    DevBuf<double> block_center, pos_build;
    DevBuf<float> block_half;
    DevBuf<int> block_minimg;
    DevBuf<int> jlist, jcount, tiles, task_start;
    DevBuf<int> flag;  // [0] rebuild requested
};

// Host-side set-up: units from the exception graph, Morton cell ranks, sizes.
void nlist_setup(NlistDev& nl, int n, const std::vector<int>& exc1, const std::vector<int>& exc2,
                 double cutoff, double skin, const Box& box, hipStream_t s);

// Requests a rebuild when an atom moved more than skin/2 since the last build
// (or always when force != 0).  Device-side: usable inside a captured graph.
void nlist_check(NlistDev& nl, const double* pos, const Box* box_dev, int force, hipStream_t s);

// Rebuilds the list if requested (every kernel exits early otherwise).
void nlist_build(NlistDev& nl, const double* pos, const Box* box_dev, unsigned* status, hipStream_t s);

}  // namespace ommhip
