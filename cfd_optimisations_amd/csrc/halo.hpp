// halo.hpp -- per-tile halo lists of the single-kernel step (built on the device).
//
// Interior cells are cut into tiles of TC consecutive cells (the block's
// cells).  For tile t with cells T:
//   H1(t)  interior cells outside T that share a face with a cell of T; the
//          block recomputes their gradient, limiter and face states;
//   S2(t)  the cells whose state is also needed and that are neither in T nor
//          in H1: first the ghosts of T's boundary faces (they receive a face
//          state), then the remaining neighbours of H1 cells (interior or ghost).
// Each list is sorted by cell id.  A block stages the states of T, H1, S2 (in
// that order: local index = position) and the face states of T, H1 and the
// ghosts of T in LDS.  The local index of every neighbour of a T or H1 cell is
// resolved once, at build time, into 16-bit tables (lt, lh), so the step
// kernel reads it with the cell's inputs instead of searching the lists.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace hcfd {

// Entry of the local-index tables lt / lh for one neighbour slot: the
// neighbour's local index among the tile's staged cells (bits 0-9), the face's
// slot in the neighbour's own list (bits 10-12), whether the neighbour is a
// ghost (bit 13) and whether the cell is the face's R side (bit 14); -1 for an
// empty slot.  With these the single-kernel step reads neither the neighbour
// planes nor the side bits: a ghost's cell id is its entry of the tile's S2 list.
constexpr int LT_LOCAL_MASK = (1 << 10) - 1;
constexpr int LT_TWIN_SHIFT = 10;
constexpr int LT_TWIN_MASK = 7;
constexpr int LT_GHOST = 1 << 13;
constexpr int LT_SIDE_R = 1 << 14;

// Cells per tile of the single-kernel step (blocks of four wavefronts): FP64
// three wavefronts (the fourth computes the tile's halo ring, about 54 cells on
// triangles), FP32-mixed all four (the ring, about 61 cells, in extra rounds).
constexpr int FUSED_CELLS = 192;
constexpr int FUSED_CELLS_MIXED = 256;
inline int fused_cells(bool mixed) { return mixed ? FUSED_CELLS_MIXED : FUSED_CELLS; }

struct HaloTables {
    int TC = 0, ntiles = 0;
    int* tile_start = nullptr;  // [ntiles + 1] first cell of each tile (tiles hold at most TC cells)
    int split_tiles = 0;        // tiles added by halving tiles whose halo exceeded the LDS budget
    int* h1_off = nullptr;  // [ntiles + 1]
    int* h1 = nullptr;      // H1 cell ids, tile by tile
    int* s_off = nullptr;   // [ntiles + 1]
    int* s_ngt = nullptr;   // [ntiles] ghosts of the tile's own cells at the head of its S2 list
    int* s2 = nullptr;      // S2 cell ids, tile by tile
    int smax = 0;           // max over tiles of |T| + |H1| + |S2| (staged states)
    int fsmax = 0;          // max over tiles of |T| + |H1| + ghosts of T (staged face states)
    int h1max = 0, s2max = 0;
    long long nh1 = 0, ns2 = 0;
    // packed entry (LT_* above) of neighbour slot k:
    short* lt = nullptr;    // [nslot][NcPad] of interior cell i in its own tile; -1 for an empty slot
    short* lh = nullptr;    // [nslot][nh1] of H1 entry e (index into h1) in the tile that lists it
    long long unresolved = 0;  // neighbours not found among the staged cells (0 unless the lists are wrong)
    // The inputs of every H1 entry, copied next to the lists in the context's
    // precision: per-slot geometry in the cell arrays' 64-entry block layout
    // (geo_index, entries padded to a multiple of 64) and 1/V.  A
    // block reads its ring's inputs as contiguous runs instead of gathering
    // them at scattered cell ids (build_ring_inputs).
    double* hgeo = nullptr;
    double* hiv = nullptr;
    float* hgeo32 = nullptr;
    float* hiv32 = nullptr;
    int2* inst_tiles = nullptr;  // [ninst] first and last tile holding a cell of each instance
};

// The lists of one tile as the kernels see them, and the local index of cell j
// among the tile's staged cells (T, then H1, then S2), or -1.
struct TileLists {
    int tb, tc, nh1, ngt, ns2, Nint;
    const int* sH1;
    const int* sS2;
};

__device__ __forceinline__ int lds_lower_bound(const int* a, int n, int x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int local_of(const TileLists& L, int j) {
    const unsigned d = (unsigned)(j - L.tb);
    if (d < (unsigned)L.tc) return (int)d;
    if (j < L.Nint) {
        const int p = lds_lower_bound(L.sH1, L.nh1, j);
        if (p < L.nh1 && L.sH1[p] == j) return L.tc + p;
    } else {
        const int p = lds_lower_bound(L.sS2, L.ngt, j);
        if (p < L.ngt && L.sS2[p] == j) return L.tc + L.nh1 + p;
    }
    const int nr = L.ns2 - L.ngt;
    const int p = lds_lower_bound(L.sS2 + L.ngt, nr, j);
    if (p < nr && L.sS2[L.ngt + p] == j) return L.tc + L.nh1 + L.ngt + p;
    return -1;
}

// Builds the tables from the neighbour planes cnb ([nslot][NcPad], neighbour |
// twin << 29 or SLOT_EMPTY) of the Nint interior cells.  Tiles start as runs of
// TC consecutive cells; a tile whose staged cells would need more LDS than
// lds_budget (the share of one block at the intended occupancy) is halved, and
// the lists rebuilt, for a few rounds.
void build_halo_tables(int Nint, int NcPad, int nslot, const int* cnb, const int* cside, int TC, bool mixed,
                       size_t lds_budget,
                       HaloTables& out, hipStream_t st);
void free_halo_tables(HaloTables& t);

// The first and last tile of each instance's interior range [inst_off[b],
// inst_off[b + 1]) into inst_tiles (batches: the CFL reduction's tile ranges).
void build_instance_tiles(const int* inst_off, int ninst, HaloTables& t, hipStream_t st);

// Copies the inputs of every H1 entry into hgeo/hiv (FP64 planes cgeo, iv) or
// hgeo32/hiv32 (FP32-mixed planes cgeo32, iv32: pass null cgeo) from the
// per-slot planes of pitch P.
void build_ring_inputs(int nslot, long long P, const double* cgeo, const double* iv, const float* cgeo32,
                       const float* iv32, HaloTables& t, hipStream_t st);

// Bytes of dynamic LDS the single-kernel step needs (block of 256 lanes):
// staged states (FP64, 32 B, in both precisions), face states (FP64: 32 B,
// FP32-mixed increments: 16 B) and the tile's CFL slots (24 B, in 32).
inline size_t fused_lds_bytes(bool mixed, int maxdeg, int smax, int fsmax) {
    const size_t per = mixed ? 16 : 32;
    return 32 * (size_t)smax + per * (size_t)maxdeg * fsmax + 32;
}

}  // namespace hcfd
