// passes.hpp -- the two fused passes of one explicit step and their launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace hcfd {

// Arguments of both passes (by value; every pointer is device memory).
struct PassArgs {
    int Nint = 0;          // interior cells [0, Nint); ghosts [Nint, Nc)
    int Nc = 0;
    int NcPad = 0;         // slot table pitch
    int ninst = 1;
    int ntiles = 0;        // ceil(Nint / tile_cells)
    int tile_cells = 0;    // cells per block
    int tiles_per_xcd = 0; // ceil(ntiles / 8) when xcd_remap
    int xcd_remap = 0;
    double gamma = 1.4;
    const int* cnb = nullptr;         // [MAXDEG][NcPad] neighbour | twin slot << 29, or SLOT_EMPTY
    const int* cside = nullptr;       // [NcPad] side bits (bit k: R side of slot k's face)
    const double* cgeo = nullptr;     // per-slot geometry, FP64, GEO_PLANES x MAXDEG planes in 64-cell blocks (geo_index)
    const float* cgeo32 = nullptr;    // same, FP32 (mixed path; the normal planes hold A n)
    const double* iv = nullptr;
    const float* iv32 = nullptr;
    const double* U = nullptr;   // current state (interior + ghosts)
    double* Unext = nullptr;     // next state
    double* FS = nullptr;        // face states, FP64: [MAXDEG][2][NcPad] double2 planes
    float* FS32 = nullptr;       // face states, FP32-mixed: [MAXDEG][NcPad] float4
    const int* inst_off = nullptr;      // [ninst+1] interior ranges
    const int* chunk_inst = nullptr;    // [ceil(Nint/64)] instance of cell 64c
    double* pmin_first = nullptr;       // [ntiles] per-tile CFL partial (maximum rate) of the tile's first instance
    double* pmin_last = nullptr;        // [ntiles] ... of its last instance
    unsigned long long* dtacc = nullptr;// [ninst + 1] maximum rate (bit pattern) of cells in neither
                                        // (slot ninst: the one-instance dt reduction's block ticket)
    const double* dt = nullptr;         // [ninst] dt of this step
    const BCEntry* bctab = nullptr;
    const int* ghost_bc = nullptr;
    int* fault = nullptr;               // [ninst]
    // single-kernel step (halo tiles, see halo.hpp)
    const int* tile_start = nullptr;    // [ntiles+1] variable tiles (single-kernel step, CFL bound), else null
    const int2* inst_tiles = nullptr;   // [ninst] first and last tile of each instance (variable tiles, batches)
    const int* h1_off = nullptr;
    const int* h1 = nullptr;
    const int* s_off = nullptr;
    const int* s_ngt = nullptr;
    const int* s2 = nullptr;
    const short* lt = nullptr;          // [nslot][NcPad] local index of each neighbour in the cell's tile
    const short* lh = nullptr;          // [nslot][nh1tot] the same for each H1 entry
    long long nh1tot = 1;
    const double* hgeo = nullptr;       // per-slot geometry of each H1 entry (geo_index layout), FP64
    const double* hiv = nullptr;        // [nh1tot]
    const float* hgeo32 = nullptr;      // the same, FP32-mixed
    const float* hiv32 = nullptr;
    int smax = 0, fsmax = 0;
    double* tacc = nullptr;             // [ninst] simulated time
    double* lastdt = nullptr;           // [ninst] dt of the latest step
    int advance = 0;                    // first stage of a step: block 0 of pass B (or of the single
                                        // kernel) adds dt to tacc and stores it in lastdt
    // Multi-stage (SSP-RK) steps: a stage after the first writes
    // rk_a * U0 + rk_b * (U + dt R), U0 = the step's starting state; null U0 = U + dt R.
    const double* U0 = nullptr;
    double rk_a = 0.0, rk_b = 1.0;
};

// Variant selectors.
struct PassVariant {
    int precision = HCFD_FP64;  // HCFD_FP64 / HCFD_MIXED
    bool cfl = false;           // pass A computes the CFL partials
    int width = 1;              // AoSoA width of U: 1, 8, 64
    int maxdeg = 3;             // slot count: 3, 4 or 8
    int lpc = 1;                // pass A lanes per cell: 1 or 4 (one lane per conserved variable)
};

// Launch the passes with `grid` x `block` threads on `st`.
hipError_t launch_pass_a(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st);
hipError_t launch_pass_b(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st);
// Per-instance CFL reduction: dt[b] = cfl * (1 / maximum rate over the instance's tile partials).
hipError_t launch_dt_reduce(const PassArgs& a, double cfl, double* dt, hipStream_t st);
// Single-kernel step over halo tiles of 256 cells, block 256, `lds` bytes of dynamic LDS.
hipError_t launch_fused_step(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st);
// CFL bound of the current state (tile partials), block 256.
hipError_t launch_cfl_bound(const PassVariant& v, const PassArgs& a, int grid, hipStream_t st);
bool pass_variant_supported(const PassVariant& v);
int pass_degree_bucket(int maxdeg);

}  // namespace hcfd
