// fused_impl.hpp -- the single-kernel step over halo tiles (included by
// fused_fp64.hip and fused_mixed.hip).
//
// One block of four wavefronts owns one tile T of consecutive interior cells
// and performs a whole step of Euler::run (src/Base/Euler.cpp:85-97) for it.
// FP64: at most FUSED_CELLS (192) cells, one per lane of three wavefronts, the
// fourth computing the tile's halo ring; FP32-mixed: at most 256 cells, one per
// lane, the last lanes taking the ring in extra rounds (fused_lanes).  Steps:
//   1. stage in LDS the states U^n of T, of its halo ring H1 and of their
//      other neighbours S2 (halo.hpp);
//   2. for every cell of T and H1: Green-Gauss gradient (GaussGreen.cpp:14-40),
//      min/max limiter (Limiter.cpp:19-77) and the MUSCL face state of each of
//      its faces (Flux.cpp:83-87), into LDS; cells of T also produce the ghost
//      side of their boundary faces (BoundaryConditions/Base.cpp:29-53);
//   3. for every cell of T: Rusanov flux of each face from the two staged face
//      states (Flux.cpp:52-109), R * 1/V, U^{n+1} = U^n + dt R
//      (Euler.cpp:107-113) [+ the SSP-RK stage combination], the ghost states
//      of its boundary faces (BoundaryConditions/{Symmetry,Wall,Inlet,Outlet}.cpp)
//      and, in CFL mode, the CFL bound of U^{n+1}: the tile partials give dt of
//      the next step.  A multi-stage step is one launch per stage; the CFL
//      bound is taken by the last stage only.
// H1 cells are computed by every tile that borders them, with the same
// arithmetic as in their own tile, so each face state equals the one the
// two-pass kernels produce (passes_impl.hpp) bit for bit; the results are those
// of the reference.  U^n is read and U^{n+1} written once per step; face states
// never leave the CU.
#pragma once
#include <type_traits>

#include "dispatch.hpp"
#include "halo.hpp"
#include "passes_impl.hpp"

namespace hcfd {

constexpr int FUSED_TILE = 256;  // threads per block: four wavefronts (tile cells: fused_cells, halo.hpp)

// LDS element types: FP64 cell state / face state = two double2; FP32 = one float4.
template <typename T>
struct FusedTypes;
template <>
struct FusedTypes<double> {
    using V = double2;
    static constexpr int PER = 2;  // 16-byte units per state
    __device__ static void put(V* s, int pitch, int l, const double u[NV]) {
        s[l] = make_double2(u[0], u[1]);
        s[pitch + l] = make_double2(u[2], u[3]);
    }
    __device__ static void get(const V* s, int pitch, int l, double u[NV]) {
        const double2 x = s[l], y = s[pitch + l];
        u[0] = x.x; u[1] = x.y; u[2] = y.x; u[3] = y.y;
    }
};
template <>
struct FusedTypes<float> {
    using V = float4;
    static constexpr int PER = 1;
    __device__ static void put(V* s, int, int l, const float u[NV]) { s[l] = make_float4(u[0], u[1], u[2], u[3]); }
    __device__ static void get(const V* s, int, int l, float u[NV]) {
        const float4 x = s[l];
        u[0] = x.x; u[1] = x.y; u[2] = x.z; u[3] = x.w;
    }
};

// Per-cell inputs of pass A and pass B, read from global memory once.
template <typename T, int MAXDEG>
struct CellIn {
    int lq[MAXDEG];   // packed local-index entry of each slot (LT_*, halo.hpp), -1: empty slot
    T A[MAXDEG], n0[MAXDEG], n1[MAXDEG], If[MAXDEG], r0[MAXDEG], r1[MAXDEG];
    T iv;
};

// Entry g of planes of pitch P: a cell's own planes (g = cell id, P = NcPad)
// or the copy of the tile's ring inputs (g = H1 entry, P = nh1tot).
template <typename T, int MAXDEG>
__device__ __forceinline__ void load_cell_in(const T* __restrict__ gp, const T* __restrict__ ivp, long long P, int g,
                                             CellIn<T, MAXDEG>& c) {
    c.iv = plane_at(ivp, g);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        // the mixed path's normal planes hold A n (A unused)
        c.A[k] = sizeof(T) == 8 ? geo<MAXDEG>(gp, P, GEO_A, k, g) : T(0);
        c.n0[k] = geo<MAXDEG>(gp, P, GEO_N0, k, g);
        c.n1[k] = geo<MAXDEG>(gp, P, GEO_N1, k, g);
        c.If[k] = geo<MAXDEG>(gp, P, GEO_IF, k, g);
        c.r0[k] = geo<MAXDEG>(gp, P, GEO_R0, k, g);
        c.r1[k] = geo<MAXDEG>(gp, P, GEO_R1, k, g);
    }
}
template <typename T, int MAXDEG>
__device__ __forceinline__ void load_cell_in(const PassArgs& a, int g, CellIn<T, MAXDEG>& c) {
    load_cell_in<T, MAXDEG>(PassATypes<T>::geo(a), PassATypes<T>::ivp(a), a.NcPad, g, c);
}
// Inputs of H1 entry e (the ring cell a.h1[e]) from the ring copy.
template <typename T, int MAXDEG>
__device__ __forceinline__ void load_ring_in(const PassArgs& a, int e, CellIn<T, MAXDEG>& c) {
    load_cell_in<T, MAXDEG>(PassATypes<T>::ring_geo(a), PassATypes<T>::ring_iv(a), a.nh1tot, e, c);
}

// Pass A of one staged cell (local index l): gradient, limiter and the face
// state of each of its faces into LDS (a tile cell also writes the ghost side
// of its boundary faces).  lnb_out receives the local neighbour indices and,
// in FP32-mixed, dnb_out the rounded differences u_nb - u of the neighbour
// states (pass B forms the face's neighbour-side increment from the same values).
// Empty slots carry zero geometry (the packing writes zeros), so every sum and
// bound is exactly that of the two-pass kernels, which skip them.
template <typename T, int MAXDEG>
__device__ __forceinline__ void fused_cell_a(const PassArgs& a, const TileLists& L, const CellIn<T, MAXDEG>& c, int l,
                                             bool own, const double2* __restrict__ sU, int su_pitch,
                                             typename FusedTypes<T>::V* __restrict__ sFS, int fs_pitch,
                                             int lnb_out[MAXDEG], T dnb_out[MAXDEG][NV], bool& bad) {
    using FT = FusedTypes<T>;
    // FP64: the state itself; FP32-mixed: differences from it (u = 0, see
    // rusanov_delta), formed from the FP64 staged states
    constexpr bool DELTA = sizeof(T) == 4;
    double ud[NV];
    FusedTypes<double>::get(sU, su_pitch, l, ud);
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = DELTA ? T(0) : (T)ud[v];
    T g2[NV][2], umin[NV], umax[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g2[v][0] = T(0);
        g2[v][1] = T(0);
        umin[v] = u[v];
        umax[v] = u[v];
    }
    // Cells of at most four faces run every slot without a branch: an empty
    // slot takes the cell itself as its neighbour, and with its zero geometry
    // it adds exact zeros to the gradient, leaves min/max as they are and gives
    // d = 0 in the limiter.  Larger cells skip their empty slots.
    constexpr bool DENSE = MAXDEG <= 4;
    int ln[MAXDEG];
    unsigned ghosts = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        const int e = c.lq[k];
        ln[k] = e;  // -1: empty slot
        if (!DENSE && e < 0) continue;
        const int q = e < 0 ? l : (e & LT_LOCAL_MASK);  // the neighbour's place among the staged cells
        double und[NV];
        FusedTypes<double>::get(sU, su_pitch, q, und);
        T un[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            un[v] = DELTA ? (T)(und[v] - ud[v]) : (T)und[v];
            dnb_out[k][v] = un[v];
        }
        const T If = c.If[k];
        const T omI = T(1) - If;
        const bool isR = (e & LT_SIDE_R) != 0;
        const T wo = isR ? omI : If, wn = isR ? If : omI;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const T uf = DELTA ? un[v] * wn : madd(u[v], wo, un[v] * wn);
            const T au = DELTA ? uf : uf * c.A[k];  // mixed: n holds A n
            g2[v][0] = madd(au, c.n0[k], g2[v][0]);
            g2[v][1] = madd(au, c.n1[k], g2[v][1]);
            umin[v] = dmin(umin[v], un[v]);
            umax[v] = dmax(umax[v], un[v]);
        }
        if (e >= 0 && (e & LT_GHOST)) ghosts |= 1u << k;
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g2[v][0] *= c.iv;
        g2[v][1] *= c.iv;
    }
    T phi[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        // an empty slot's r is zero, so its d = 0 changes neither bound
        T dpos = T(0), dneg = T(0);
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            if (!DENSE && ln[k] < 0) continue;
            const T d = dot_r(g2[v][0], g2[v][1], c.r0[k], c.r1[k]);
            dpos = dmax(dpos, d);
            dneg = dmin(dneg, d);
        }
        phi[v] = limiter_phi(dpos, dneg, umin[v] - u[v], umax[v] - u[v]);
    }
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (DENSE || ln[k] >= 0) {  // an empty slot's face state is never read
            T d[NV], uf[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                d[v] = dot_r(g2[v][0], g2[v][1], c.r0[k], c.r1[k]);
                uf[v] = madd(phi[v], d[v], u[v]);
            }
            FT::put(sFS + (size_t)(k * FT::PER) * fs_pitch, fs_pitch, l, uf);
            // mixed: the ghost's increment is this cell's (pass B uses it directly)
            if (!DELTA && own && ((ghosts >> k) & 1u)) {
                // ghost side of a boundary face: U_g + phi_L ((-G_L) . rR), rR = -rL;
                // the ghost's slot lies in the face-state region [tc + nh1, fs_pitch)
                const int gq = ln[k] & LT_LOCAL_MASK;
                if (gq < L.tc + L.nh1 || gq >= fs_pitch) {
                    bad = true;
                    continue;
                }
                double ug[NV];
                T ufg[NV];
                FusedTypes<double>::get(sU, su_pitch, gq, ug);
#pragma unroll
                for (int v = 0; v < NV; v++) ufg[v] = madd(phi[v], d[v], (T)ug[v]);
                FT::put(sFS, fs_pitch, gq, ufg);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) lnb_out[k] = ln[k];
}

// CFL rate iv sum_f A_f (|u.n_f| + c) of an FP64 state, from a cell's inputs
// (the cell's bound is 1/rate; see cfl_rate); bit k of `present`: slot k holds a face.
template <typename T, int MAXDEG>
__device__ __forceinline__ double cfl_bound_in(double gamma, const CellIn<T, MAXDEG>& ci, const double u[NV],
                                               unsigned present) {
    const double gmo = gamma - 1;
    const double ir = 1.0 / u[0];
    const double ux = u[1] * ir, uy = u[2] * ir;
    const double p = gmo * (u[3] - 0.5 * (u[1] * ux + u[2] * uy));
    const double c = sqrt(gamma * p * ir);
    double lam = 0.0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (!((present >> k) & 1u)) continue;
        lam += cfl_face(ux, uy, c, ci.A[k], ci.n0[k], ci.n1[k]);
    }
    return cfl_rate((double)ci.iv * lam);
}

// The neighbours' local indices of a staged cell: entry e of the planes of
// `loc` (pitch lpitch), the cell's own entry of lt for a tile cell, its H1
// entry of lh for a ring cell.
template <typename T, int MAXDEG>
__device__ __forceinline__ void load_local(const short* __restrict__ loc, long long lpitch, int e,
                                           CellIn<T, MAXDEG>& c) {
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) c.lq[k] = plane_at(loc + k * lpitch, e);
}

template <typename T, int MAXDEG>
__device__ __forceinline__ double cfl_bound(const PassArgs& a, int g, const double u[NV]) {
    CellIn<T, MAXDEG> ci;
    load_cell_in<T, MAXDEG>(a, g, ci);
    unsigned present = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++)
        if (cnb_at(a, k, g) != SLOT_EMPTY) present |= 1u << k;
    return cfl_bound_in<T, MAXDEG>(a.gamma, ci, u, present);
}

// CFL partials of one tile without a block barrier: each wavefront folds its
// maximum rate into LDS with an atomic, and the last of the tile's wavefronts to
// arrive (LDS counter) publishes the tile partial and resets the slots.  The
// rates are non-negative (or +inf; cfl_rate maps NaN to 0), so the bit patterns
// order like the values.  For batches the
// tile's first and last instance get their own slot; cells of instances
// strictly inside the tile use global atomics, as in cfl_tile.
struct TileMin {
    unsigned long long first, last;
    unsigned arrived;
};

// inst: the lane's cell's instance (-1 without a cell); bf, bl: the instances of
// the tile's first and last cell.  A tile inside one instance (bf == bl, the
// common case) needs only the first partial.
// NWAVES: the wavefronts that arrive (those that carry tile cells).
template <bool MULTI, int NWAVES>
__device__ __forceinline__ void cfl_tile_lds(const PassArgs& a, int tile, int inst, int bf, int bl, double rate,
                                             TileMin* tm) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int nwaves = NWAVES;
    double vf = rate, vl = 0.0;
    const bool two = MULTI && bf != bl;  // block-uniform
    if (two) {
        vf = 0.0;
        if (inst >= 0) {
            if (inst == bf) vf = rate;
            else if (inst == bl) vl = rate;
            else atomicMax(&a.dtacc[inst], dbits(rate));
        }
        vl = wave_max_dpp(vl);
    }
    vf = wave_max_dpp(vf);
    if (lane == 0) {
        atomicMax(&tm->first, dbits(vf));
        if (two) atomicMax(&tm->last, dbits(vl));
        __threadfence_block();
        if (atomicAdd(&tm->arrived, 1u) == nwaves - 1) {
            __threadfence_block();
            a.pmin_first[tile] = __longlong_as_double((long long)atomicExch(&tm->first, 0ull));
            if constexpr (MULTI) a.pmin_last[tile] = __longlong_as_double((long long)atomicExch(&tm->last, 0ull));
            tm->arrived = 0;
        }
    }
}

// Minimum waves per SIMD for the single-kernel step on triangles (bounds its
// VGPRs; the LDS of a tile allows HCFD_FUSED_FP64_BLOCKS FP64 blocks of four
// waves per CU and HCFD_FUSED_MIXED_BLOCKS FP32-mixed ones; cells of more faces
// two).  FP64 at four waves: 122 VGPRs without spills, the ring wavefront
// holding no tile cell's inputs.  FP32-mixed at five: 89.
#ifndef HCFD_WAVES_FUSED
#define HCFD_WAVES_FUSED 4
#endif
#ifndef HCFD_WAVES_FUSED_MIXED
#define HCFD_WAVES_FUSED_MIXED 5
#endif

// Lane roles of a block.  RW (FP64): three wavefronts carry the tile's (at most
// FUSED_CELLS) cells and one computes pass A of the tile's halo ring, so the
// ring's inputs load during the staging and no lane holds the inputs of a tile
// cell and a ring cell at once; the ring's wavefront is rotated from tile to
// tile, so every SIMD gets ring work (pass A only) as well as tile work.
// Otherwise (FP32-mixed, where the extra pass-B-free wavefront costs more issue
// than it saves): every lane carries a tile cell and the lanes from the top of
// a rotated order take the ring cells in extra rounds.
struct FusedLanes {
    bool ring;  // RW: this lane's wavefront computes the halo ring
    int tl;     // the lane's tile cell (RW ring wavefront: its ring lane 0..63)
    int rq;     // !RW: the lane's first ring entry (none when >= the ring's size)
};
template <bool RW>
__device__ __forceinline__ FusedLanes fused_lanes(int tile) {
    const int tid = threadIdx.x, w = tid / WAVE, lane = tid & (WAVE - 1);
    const int rw = (int)((blockIdx.x + tile) & (FUSED_TILE / WAVE - 1));
    FusedLanes f;
    if constexpr (RW) {
        f.ring = w == rw;
        f.tl = f.ring ? lane : (w < rw ? w : w - 1) * WAVE + lane;
        f.rq = FUSED_TILE;
    } else {
        f.ring = false;
        f.tl = tid;
        f.rq = FUSED_TILE - 1 - ((tid + WAVE * rw) & (FUSED_TILE - 1));
    }
    return f;
}

// LOOP: a block walks several tiles (a persistent grid); otherwise each block
// owns at most one tile, and the kernel holds no loop, so its register budget
// serves that one tile only (the loop's variant needs more registers).
template <typename T, int W, int MAXDEG, bool CFL, bool MULTI, bool LOOP>
__global__ __launch_bounds__(FUSED_TILE) __attribute__((amdgpu_waves_per_eu(
    MAXDEG > 3 ? 2 : (sizeof(T) == 4 ? HCFD_WAVES_FUSED_MIXED : HCFD_WAVES_FUSED)))) void
k_fused_step(
    PassArgs a) {
    using FT = FusedTypes<T>;
    using V = typename FT::V;
    extern __shared__ double2 fused_lds[];
    const int su_pitch = a.smax, fs_pitch = a.fsmax;
    // staged states are FP64 in both precisions (the mixed path forms its
    // differences from them); face states (mixed: increments) are T
    double2* sU = fused_lds;
    V* sFS = reinterpret_cast<V*>(sU + (size_t)2 * su_pitch);
    TileMin* tmin = reinterpret_cast<TileMin*>(sFS + (size_t)FT::PER * MAXDEG * fs_pitch);
    const int tid = threadIdx.x;
    if (tid == 0) {
        tmin->first = 0ull;
        tmin->last = 0ull;
        tmin->arrived = 0;
    }  // first read after the tile loop's first barrier

    constexpr bool RW = sizeof(T) == 8;  // a ring wavefront (fused_lanes)
    advance_clocks(a);  // dt of this step: fixed, or reduced from the CFL bound of U^n
    int tfirst, step, tend;
    tile_walk(a, tfirst, step, tend);
    auto run_tile = [&](int tile) {
        TileLists L;
        L.tb = a.tile_start[tile];
        L.tc = a.tile_start[tile + 1] - L.tb;
        const int h0 = a.h1_off[tile], s0 = a.s_off[tile];
        L.nh1 = a.h1_off[tile + 1] - h0;
        L.ns2 = a.s_off[tile + 1] - s0;
        L.ngt = a.s_ngt[tile];
        L.Nint = a.Nint;
        L.sH1 = nullptr;  // neighbours' local indices come from lt / lh
        L.sS2 = nullptr;
        const FusedLanes fl = fused_lanes<RW>(tile);
        const int tl = fl.tl;
        const int i = L.tb + tl;
        const bool hasT = !fl.ring && tl < L.tc;
        // instances of the tile's first and last cell (batches)
        int bf = 0, bl = 0;
        if constexpr (MULTI) {
            bf = instance_from(a.inst_off, a.chunk_inst, L.tb);
            bl = instance_from(a.inst_off, a.chunk_inst, L.tb + L.tc - 1);
        }
        // the lane's tile cell, or the ring wavefront's first ring cell: its
        // inputs are in flight while the tile is staged; without a ring
        // wavefront, so are those of the first ring cell of the lanes that take one
        CellIn<T, MAXDEG> ci, ch;
        double uown[NV] = {0, 0, 0, 0};  // U^n of the lane's tile cell (read back from LDS in pass B)
        if (hasT) {
            load_cell_in<T, MAXDEG>(a, i, ci);
            load_local(a.lt, a.NcPad, i, ci);
            load_cell<W, NV>(a.U, i, uown);
        } else if (RW && fl.ring && tl < L.nh1) {
            load_ring_in<T, MAXDEG>(a, h0 + tl, ci);
            load_local(a.lh, a.nh1tot, h0 + tl, ci);
        }
        if (!RW && fl.rq < L.nh1) {
            load_ring_in<T, MAXDEG>(a, h0 + fl.rq, ch);
            load_local(a.lh, a.nh1tot, h0 + fl.rq, ch);
        }
        // staged halo cell q = tid (ring H1, then S2): its id and state are in
        // flight before the barrier too; further rounds (more than 256 staged
        // halo cells) load after it
        const int nhalo = L.nh1 + L.ns2;
        int gq = 0;
        double uq[NV] = {0, 0, 0, 0};
        if (tid < nhalo) {
            gq = tid < L.nh1 ? a.h1[h0 + tid] : a.s2[s0 + tid - L.nh1];
            load_cell<W, NV>(a.U, gq, uq);
        }
        __syncthreads();  // the previous tile's readers are done
        if (hasT) FusedTypes<double>::put(sU, su_pitch, tl, uown);
        for (int q = tid; q < nhalo; q += FUSED_TILE) {
            int g = gq;
            double ud[NV];
            if (q == tid) {
#pragma unroll
                for (int v = 0; v < NV; v++) ud[v] = uq[v];
            } else {
                g = q < L.nh1 ? a.h1[h0 + q] : a.s2[s0 + q - L.nh1];
                load_cell<W, NV>(a.U, g, ud);
            }
            FusedTypes<double>::put(sU, su_pitch, L.tc + q, ud);
        }
        __syncthreads();
        // pass A: the tile cells and the halo ring
        bool bad = false;
        int lnb[MAXDEG];
        T dnb[MAXDEG][NV];  // FP32-mixed: u_nb - u of each slot, kept for pass B
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            lnb[k] = -1;
#pragma unroll
            for (int v = 0; v < NV; v++) dnb[k][v] = T(0);
        }
        if (hasT) fused_cell_a<T, MAXDEG>(a, L, ci, tl, true, sU, su_pitch, sFS, fs_pitch, lnb, dnb, bad);
        if constexpr (RW) {
            if (fl.ring) {
                for (int q = tl; q < L.nh1; q += WAVE) {
                    if (q != tl) {
                        load_ring_in<T, MAXDEG>(a, h0 + q, ci);
                        load_local(a.lh, a.nh1tot, h0 + q, ci);
                    }
                    int ln[MAXDEG];
                    T dn[MAXDEG][NV];
                    fused_cell_a<T, MAXDEG>(a, L, ci, L.tc + q, false, sU, su_pitch, sFS, fs_pitch, ln, dn, bad);
                }
            }
        } else {
            for (int q = fl.rq; q < L.nh1; q += FUSED_TILE) {
                if (q != fl.rq) {
                    load_ring_in<T, MAXDEG>(a, h0 + q, ch);
                    load_local(a.lh, a.nh1tot, h0 + q, ch);
                }
                int ln[MAXDEG];
                T dn[MAXDEG][NV];
                fused_cell_a<T, MAXDEG>(a, L, ch, L.tc + q, false, sU, su_pitch, sFS, fs_pitch, ln, dn, bad);
            }
        }
        __syncthreads();
        if (RW && fl.ring) {  // no pass B on the ring wavefront
            if (bad) a.fault[0] = 1;
            return;
        }
        // pass B over the tile
        double rate = 0.0;
        int lane_inst = -1;
        if (hasT) {
            const long long P = a.NcPad;
            int inst = bf;
            if (MULTI && bf != bl)
                while (a.inst_off[inst + 1] <= i) inst++;
            lane_inst = inst;
            const double dt = a.dt[inst];
            double R[NV] = {0.0, 0.0, 0.0, 0.0};
            // U^n of the cell back from LDS (not held in registers through pass A)
            FusedTypes<double>::get(sU, su_pitch, tl, uown);
            FluxBase fb;
            if constexpr (sizeof(T) == 4) fb = flux_base((float)a.gamma, uown);
            unsigned gmask = 0, present = 0;
#pragma unroll
            for (int k = 0; k < MAXDEG; k++) {
                const int e = lnb[k];  // packed local-index entry (LT_*)
                if (e < 0) continue;   // empty slot
                present |= 1u << k;
                const int tw = (e >> LT_TWIN_SHIFT) & LT_TWIN_MASK;
                const bool gnb = (e & LT_GHOST) != 0;
                int q = e & LT_LOCAL_MASK;
                if (q >= fs_pitch || tw >= MAXDEG) {
                    bad = true;
                    q = tl;
                }
                T Uo[NV], Un[NV];
                FT::get(sFS + (size_t)(k * FT::PER) * fs_pitch, fs_pitch, tl, Uo);
                T F[NV];
                if constexpr (sizeof(T) == 4) {
                    // increments of the two face states (see pass_b_cell_mixed);
                    // a ghost's increment is this cell's own, read again from the
                    // own slot (one LDS read either way, no branch per face); the
                    // slot offsets are 24-bit products (slot * PER < 16, fs_pitch <= 1024)
                    const int ks = gnb ? k : (tw < MAXDEG ? tw : 0);
                    FT::get(sFS + __umul24((unsigned)(ks * FT::PER), (unsigned)fs_pitch), fs_pitch, gnb ? tl : q, Un);
#pragma unroll
                    for (int v = 0; v < NV; v++) Un[v] += dnb[k][v];  // (float)(u_nb - u), as pass A rounded it
                    rusanov_delta((float)a.gamma, fb, Uo, Un, ci.n0[k], ci.n1[k], F);
                } else {
                    FT::get(sFS + __umul24((unsigned)((tw < MAXDEG ? tw : 0) * FT::PER), (unsigned)fs_pitch), fs_pitch,
                            q, Un);
                    rusanov((T)a.gamma, Uo, Un, ci.n0[k], ci.n1[k], F);
                }
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    if constexpr (sizeof(T) == 8) R[v] = R[v] - ci.A[k] * F[v];
                    else R[v] = R[v] - (double)F[v];  // rusanov_delta returns 2 A (F - E(u).n)
                }
                if (gnb) gmask |= 1u << k;
            }
            double unew[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                if constexpr (sizeof(T) == 4) R[v] *= 0.5;  // the Rusanov 1/2 of the mixed terms
                R[v] = R[v] * (double)ci.iv;
                unew[v] = uown[v] + dt * R[v];
            }
            if (a.U0) rk_combine<W>(a, i, unew);
            store_cell<W, NV>(a.Unext, i, unew);
            if (nonphysical(a.gamma, unew) || bad) a.fault[inst] = 1;
            // behind one test of the whole mask, so that the ghost states'
            // common terms (the divide of the subsonic BCs) are not formed on
            // cells without a boundary face
            if (gmask != 0u) {
#pragma unroll
                for (int k = 0; k < MAXDEG; k++) {
                    if (!((gmask >> k) & 1u)) continue;
                    // the ghost's cell id: its entry among the tile's ghosts at the
                    // head of S2 (staged right after the halo ring)
                    const int gp = (lnb[k] & LT_LOCAL_MASK) - L.tc - L.nh1;
                    if (gp < 0 || gp >= L.ngt) {
                        a.fault[inst] = 1;
                        continue;
                    }
                    const int gc = a.s2[s0 + gp];
                    // boundary faces have the interior cell on the L side: n_out = n (FP64)
                    const BCEntry e = a.bctab[a.ghost_bc[gc - a.Nint]];
                    double ug[NV];
                    ghost_state(e, a.gamma, unew, geo<MAXDEG>(a.cgeo, P, GEO_N0, k, i),
                                geo<MAXDEG>(a.cgeo, P, GEO_N1, k, i), ug);
                    store_cell<W, NV>(a.Unext, gc, ug);
                }
            }
            if constexpr (CFL) rate = cfl_bound_in<T, MAXDEG>(a.gamma, ci, unew, present);
        } else if (bad) {
            a.fault[0] = 1;
        }
        if constexpr (CFL) cfl_tile_lds<MULTI, (RW ? 3 : 4)>(a, tile, lane_inst, bf, bl, rate, tmin);
    };
    // One tile per block (the default grid) runs outside any loop: inside one,
    // the compiler hoists the tile cell's LDS addresses out of it and, in FP64,
    // spills them (measured 36-76 B per lane and 20% slower).
    if constexpr (LOOP) {
        for (int tile = tfirst; tile < tend; tile += step) run_tile(tile);
    } else {
        if (tfirst + step < tend) {  // a grid of fewer blocks than tiles: launch_fused_step_t picks LOOP
            if (tid == 0) a.fault[0] = 1;
            return;
        }
        if (tfirst < tend) run_tile(tfirst);
    }
}

// Launch of the single-kernel step, one tile per block or (LOOP) a persistent
// grid; explicitly instantiated in fused_{fp64,mixed}{,_loop}.hip so the four
// variant sets compile in parallel.
template <typename T, bool LOOP>
hipError_t launch_fused_step_t(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st) {
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(v.cfl, [&](auto CF) {
                return with_bool(a.ninst > 1, [&](auto MU) {
                    auto k = k_fused_step<T, decltype(W)::value, decltype(MD)::value, decltype(CF)::value,
                                          decltype(MU)::value, LOOP>;
                    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)lds);
                    if (e != hipSuccess) return e;
                    hipLaunchKernelGGL(k, dim3(grid), dim3(FUSED_TILE), lds, st, a);
                    return hipGetLastError();
                });
            });
        });
    });
}

// True when `grid` blocks leave some block more than one tile (tile_walk).
inline bool fused_grid_loops(const PassArgs& a, int grid) {
    const long long one_each = a.xcd_remap ? (long long)NXCD * a.tiles_per_xcd : (long long)a.ntiles;
    return (long long)grid < one_each;
}

// CFL bound of the current state U (tile partials only), before the first
// step of a run: the single-kernel step reduces the bound of U^{n+1} itself.
template <typename T, int W, int MAXDEG, bool MULTI>
__global__ __launch_bounds__(FUSED_TILE) void k_cfl_bound(PassArgs a) {
    __shared__ double sh[FUSED_TILE / WAVE];
    int tfirst, step, tend;
    tile_walk(a, tfirst, step, tend);
    for (int tile = tfirst; tile < tend; tile += step) {
        const int tb = a.tile_start[tile], tc = a.tile_start[tile + 1] - tb;
        const int i = tb + threadIdx.x;
        double rate = 0.0;
        if ((int)threadIdx.x < tc) {
            double u[NV];
            load_cell<W, NV>(a.U, i, u);
            rate = cfl_bound<T, MAXDEG>(a, i, u);
        }
        cfl_tile_range<MULTI>(a, tile, tb, tb + tc - 1, (int)threadIdx.x < tc ? i : a.Nint, rate, sh);
    }
}

}  // namespace hcfd
