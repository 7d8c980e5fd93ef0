// passes_impl.hpp -- kernel bodies of the two fused passes (included by
// passes_fp64.hip and passes_mixed.hip).
//
// One explicit step of Euler::run (src/Base/Euler.cpp:85-97) needs every cell's
// gradient and limiter before any face flux, so a step is two passes over the
// interior cells, each cell owning its own update (owner-computes gather, no
// atomics, deterministic):
//
//   pass A  gradient (GaussGreen.cpp:14-40) + limiter (Limiter.cpp:19-77)
//           [+ per-cell CFL bound]; writes phi*G (or G and phi)
//   pass B  MUSCL + Rusanov flux of every face of the cell (Flux.cpp:9-109),
//           residual * 1/V, U += dt R (Euler.cpp:107-113), and the ghost
//           states of the cell's boundary faces (BoundaryConditions/*.cpp)
//
// A face is evaluated by both of its cells, in each cell's ascending original
// face order, so every cell accumulates exactly the terms the reference adds to
// it, in the same order.  Ghost cells need no pass-A values: their gradient is
// -G_L and their limiter phi_L (BoundaryConditions/Base.cpp:29-53), both the
// interior cell's own.
#pragma once
#include "device_math.hpp"
#include "passes.hpp"

namespace hcfd {

// Tile walk: without remapping, block b walks tiles b, b+G, ...; with XCD
// remapping (hardware dispatches block b to XCD b % 8), the blocks of one XCD
// walk one contiguous eighth of the tiles so neighbour gathers hit that XCD's L2.
__device__ __forceinline__ void tile_walk(const PassArgs& a, int& t0, int& step, int& tend) {
    const int b = blockIdx.x, G = gridDim.x;
    if (a.xcd_remap) {
        const int xcd = b % NXCD, k = b / NXCD;
        const int per = a.tiles_per_xcd;
        t0 = xcd * per + k;
        step = G / NXCD;
        tend = min((xcd + 1) * per, a.ntiles);
    } else {
        t0 = b;
        step = G;
        tend = a.ntiles;
    }
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// CFL partials of one tile: the tile's first and last instance get block
// minima; cells of any instance strictly inside the tile use atomics.
template <bool MULTI>
__device__ __forceinline__ void cfl_tile(const PassArgs& a, int tile, int i, double dti, double* sh) {
    if constexpr (!MULTI) {
        const double m = block_min(dti, sh);
        if (threadIdx.x == 0) a.pmin_first[tile] = m;
    } else {
        const int c0 = tile * a.tile_cells;
        const int c1 = min(c0 + a.tile_cells, a.Nint) - 1;
        const int bf = instance_of(a.inst_off, a.ninst, c0);
        const int bl = instance_of(a.inst_off, a.ninst, c1);
        double vf = INFINITY, vl = INFINITY;
        if (i < a.Nint) {
            const int bi = instance_of(a.inst_off, a.ninst, i);
            if (bi == bf) vf = dti;
            else if (bi == bl) vl = dti;
            else atomicMin(&a.dtacc[bi], dbits(dti));
        }
        const double mf = block_min(vf, sh);
        const double ml = block_min(vl, sh);
        if (threadIdx.x == 0) {
            a.pmin_first[tile] = mf;
            a.pmin_last[tile] = ml;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 pass A
// ---------------------------------------------------------------------------
template <int W, int MAXDEG, bool EXACT, bool CFL>
__device__ __forceinline__ double pass_a_cell_fp64(const PassArgs& a, int i) {
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const double iv = a.iv[i];
    double g[NV][2], umin[NV], umax[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] = 0.0;
        g[v][1] = 0.0;
        umin[v] = u[v];
        umax[v] = u[v];
    }
    double ux = 0, uy = 0, c = 0, lam = 0.0;
    if constexpr (CFL) {
        const double gmo = a.gamma - 1;
        const double ir = 1.0 / u[0];
        ux = u[1] * ir;
        uy = u[2] * ir;
        const double p = gmo * (u[3] - 0.5 * (u[1] * ux + u[2] * uy));
        c = sqrt(a.gamma * p * ir);
    }
    double r0[MAXDEG], r1[MAXDEG];
    int deg = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        r0[k] = 0.0;
        r1[k] = 0.0;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        deg = k + 1;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
        double un[NV];
        load_cell<W, NV>(a.U, s.nb, un);
        const double If = fr.If, omI = 1 - If;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            // uf = U_L If + U_R (1 - If); G_L += uf A n, G_R -= uf A n
            const double uf = isR ? (un[v] * If + u[v] * omI) : (u[v] * If + un[v] * omI);
            const double au = uf * fr.A;
            const double x0 = au * fr.n0, x1 = au * fr.n1;
            if (isR) {
                g[v][0] = g[v][0] - x0;
                g[v][1] = g[v][1] - x1;
            } else {
                g[v][0] = g[v][0] + x0;
                g[v][1] = g[v][1] + x1;
            }
            umin[v] = fmin(umin[v], un[v]);
            umax[v] = fmax(umax[v], un[v]);
        }
        r0[k] = isR ? fr.rR0 : fr.rL0;
        r1[k] = isR ? fr.rR1 : fr.rL1;
        if constexpr (CFL) lam += fr.A * (fabs(ux * fr.n0 + uy * fr.n1) + c);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] *= iv;
        g[v][1] *= iv;
    }
    double phi[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        phi[v] = 1.0;
        const double dmn = umin[v] - u[v], dmx = umax[v] - u[v];
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            if (k < deg) {
                const double d = g[v][0] * r0[k] + g[v][1] * r1[k];
                phi[v] = fmin(phi[v], limiter_face(d, dmn, dmx));
            }
        }
    }
    if constexpr (EXACT) {
        double gg[2 * NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            gg[2 * v] = g[v][0];
            gg[2 * v + 1] = g[v][1];
        }
        store_cell<W, 2 * NV>(a.G, i, gg);
        store_cell<W, NV>(a.PHI, i, phi);
    } else {
        double pg[2 * NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            pg[2 * v] = phi[v] * g[v][0];
            pg[2 * v + 1] = phi[v] * g[v][1];
        }
        store_cell<W, 2 * NV>(a.G, i, pg);
    }
    if constexpr (CFL) return 1.0 / (iv * lam);
    return 0.0;
}

template <int W, int MAXDEG, bool EXACT, bool CFL, bool MULTI>
__global__ __launch_bounds__(256) void k_pass_a_fp64(PassArgs a) {
    __shared__ double sh[256 / WAVE];
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + threadIdx.x;
        double dti = INFINITY;
        if (i < a.Nint) dti = pass_a_cell_fp64<W, MAXDEG, EXACT, CFL>(a, i);
        if constexpr (CFL) cfl_tile<MULTI>(a, tile, i, dti, sh);
    }
}

// ---------------------------------------------------------------------------
// Ghost states after the update: BoundaryCondition::update_cells of
// Symmetry.cpp:5-24, Wall.cpp:4-15, Inlet.cpp:4-31, Outlet.cpp:4-13 and the
// corrected SubOutlet (CAoS Outlet.cpp:15-26), from the new interior state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ghost_state(const BCEntry& e, double gma, const double ul[NV], double n0, double n1,
                                            double ug[NV]) {
    switch (e.type) {
        case HCFD_SYMMETRY: {
            const double nor = ul[1] * n0 + ul[2] * n1;
            ug[0] = ul[0];
            ug[1] = ul[1] - 2 * nor * n0;
            ug[2] = ul[2] - 2 * nor * n1;
            ug[3] = ul[3];
        } break;
        case HCFD_WALL:
            ug[0] = ul[0]; ug[1] = -ul[1]; ug[2] = -ul[2]; ug[3] = ul[3];
            break;
        case HCFD_SUPINLET:
            ug[0] = e.UBC[0]; ug[1] = e.UBC[1]; ug[2] = e.UBC[2]; ug[3] = e.UBC[3];
            break;
        case HCFD_SUBINLET: {
            const double PL = (gma - 1) * (ul[3] - 0.5 * (ul[1] * ul[1] + ul[2] * ul[2]) / ul[0]);
            ug[0] = e.UBC[0]; ug[1] = e.UBC[1]; ug[2] = e.UBC[2];
            ug[3] = PL / (gma - 1) + 0.5 * e.QBC[0] * (e.QBC[1] * e.QBC[1] + e.QBC[2] * e.QBC[2]);
        } break;
        case HCFD_SUBOUTLET:
            ug[0] = ul[0]; ug[1] = ul[1]; ug[2] = ul[2];
            ug[3] = e.QBC[3] / (gma - 1) + 0.5 * (ul[1] * ul[1] + ul[2] * ul[2]) / ul[0];
            break;
        case HCFD_SUPOUTLET:
        default:
            ug[0] = ul[0]; ug[1] = ul[1]; ug[2] = ul[2]; ug[3] = ul[3];
            break;
    }
}

__device__ __forceinline__ bool nonphysical(double gma, const double u[NV]) {
    const double p = (gma - 1) * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0]);
    return !(u[0] > 0.0) || !(p > 0.0);
}

// ---------------------------------------------------------------------------
// FP64 pass B
// ---------------------------------------------------------------------------
template <int W, int MAXDEG, bool EXACT>
__device__ __forceinline__ void pass_b_cell_fp64(const PassArgs& a, int i, double dt, int inst) {
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const double iv = a.iv[i];
    double gi[2 * NV], phii[NV];
    load_cell<W, 2 * NV>(a.G, i, gi);
    if constexpr (EXACT) load_cell<W, NV>(a.PHI, i, phii);
    double R[NV] = {0.0, 0.0, 0.0, 0.0};
    int gnb[MAXDEG];
    double gn0[MAXDEG], gn1[MAXDEG];
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gnb[k] = -1;
        gn0[k] = 0.0;
        gn1[k] = 0.0;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
        const int j = s.nb;
        const bool ghost = j >= a.Nint;
        double un[NV];
        load_cell<W, NV>(a.U, j, un);
        double gj[2 * NV], phij[NV];
        if (ghost) {
#pragma unroll
            for (int q = 0; q < 2 * NV; q++) gj[q] = -gi[q];
            if constexpr (EXACT) {
#pragma unroll
                for (int v = 0; v < NV; v++) phij[v] = phii[v];
            }
            gnb[k] = j;
            gn0[k] = fr.n0;
            gn1[k] = fr.n1;
        } else {
            load_cell<W, 2 * NV>(a.G, j, gj);
            if constexpr (EXACT) load_cell<W, NV>(a.PHI, j, phij);
        }
        const double ro0 = isR ? fr.rR0 : fr.rL0, ro1 = isR ? fr.rR1 : fr.rL1;
        const double rn0 = isR ? fr.rL0 : fr.rR0, rn1 = isR ? fr.rL1 : fr.rR1;
        double uo[NV], unb[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if constexpr (EXACT) {
                uo[v] = u[v] + phii[v] * (gi[2 * v] * ro0 + gi[2 * v + 1] * ro1);
                unb[v] = un[v] + phij[v] * (gj[2 * v] * rn0 + gj[2 * v + 1] * rn1);
            } else {
                uo[v] = u[v] + (gi[2 * v] * ro0 + gi[2 * v + 1] * ro1);
                unb[v] = un[v] + (gj[2 * v] * rn0 + gj[2 * v + 1] * rn1);
            }
        }
        double F[NV];
        if (isR) rusanov(a.gamma, unb, uo, fr.n0, fr.n1, F);
        else rusanov(a.gamma, uo, unb, fr.n0, fr.n1, F);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (isR) R[v] = R[v] + fr.A * F[v];
            else R[v] = R[v] - fr.A * F[v];
        }
    }
    double unew[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        R[v] = R[v] * iv;
        unew[v] = u[v] + dt * R[v];
    }
    store_cell<W, NV>(a.Unext, i, unew);
    if (nonphysical(a.gamma, unew)) a.fault[inst] = 1;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (gnb[k] >= 0) {
            const BCEntry e = a.bctab[a.ghost_bc[gnb[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, gn0[k], gn1[k], ug);
            store_cell<W, NV>(a.Unext, gnb[k], ug);
        }
    }
}

template <int W, int MAXDEG, bool EXACT, bool MULTI>
__global__ __launch_bounds__(256) void k_pass_b_fp64(PassArgs a) {
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + threadIdx.x;
        if (i >= a.Nint) continue;
        const int inst = MULTI ? instance_of(a.inst_off, a.ninst, i) : 0;
        pass_b_cell_fp64<W, MAXDEG, EXACT>(a, i, a.dt[inst], inst);
    }
}

// ---------------------------------------------------------------------------
// FP32-mixed passes: state, gradient, limiter and flux in FP32; the residual
// summed with Kahan compensation; CFL bound, update and ghosts in FP64.
// ---------------------------------------------------------------------------
struct Kahan {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <int W, int MAXDEG, bool CFL>
__device__ __forceinline__ double pass_a_cell_mixed(const PassArgs& a, int i) {
    double ud[NV];
    load_cell<W, NV>(a.U, i, ud);
    float u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = (float)ud[v];
    const float iv = a.iv32[i];
    float g[NV][2], umin[NV], umax[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] = 0.0f;
        g[v][1] = 0.0f;
        umin[v] = u[v];
        umax[v] = u[v];
    }
    double ux = 0, uy = 0, c = 0, lam = 0.0;
    if constexpr (CFL) {
        const double gmo = a.gamma - 1;
        const double ir = 1.0 / ud[0];
        ux = ud[1] * ir;
        uy = ud[2] * ir;
        const double p = gmo * (ud[3] - 0.5 * (ud[1] * ux + ud[2] * uy));
        c = sqrt(a.gamma * p * ir);
    }
    float r0[MAXDEG], r1[MAXDEG];
    int deg = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        r0[k] = 0.0f;
        r1[k] = 0.0f;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        deg = k + 1;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec32 fr = a.faces32[s.f & (SLOT_SIDE_R - 1)];
        double und[NV];
        load_cell<W, NV>(a.U, s.nb, und);
        const float If = fr.If, omI = 1.0f - If;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const float un = (float)und[v];
            const float uf = isR ? (un * If + u[v] * omI) : (u[v] * If + un * omI);
            const float au = uf * fr.A;
            const float x0 = au * fr.n0, x1 = au * fr.n1;
            if (isR) {
                g[v][0] -= x0;
                g[v][1] -= x1;
            } else {
                g[v][0] += x0;
                g[v][1] += x1;
            }
            umin[v] = fminf(umin[v], un);
            umax[v] = fmaxf(umax[v], un);
        }
        r0[k] = isR ? fr.rR0 : fr.rL0;
        r1[k] = isR ? fr.rR1 : fr.rL1;
        if constexpr (CFL) lam += (double)fr.A * (fabs(ux * (double)fr.n0 + uy * (double)fr.n1) + c);
    }
    float pg[2 * NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] *= iv;
        g[v][1] *= iv;
        float phi = 1.0f;
        const float dmn = umin[v] - u[v], dmx = umax[v] - u[v];
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            if (k < deg) {
                const float d = g[v][0] * r0[k] + g[v][1] * r1[k];
                phi = fminf(phi, limiter_face(d, dmn, dmx));
            }
        }
        pg[2 * v] = phi * g[v][0];
        pg[2 * v + 1] = phi * g[v][1];
    }
    store_cell<W, 2 * NV>(a.G32, i, pg);
    if constexpr (CFL) return 1.0 / ((double)iv * lam);
    return 0.0;
}

template <int W, int MAXDEG, bool CFL, bool MULTI>
__global__ __launch_bounds__(256) void k_pass_a_mixed(PassArgs a) {
    __shared__ double sh[256 / WAVE];
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + threadIdx.x;
        double dti = INFINITY;
        if (i < a.Nint) dti = pass_a_cell_mixed<W, MAXDEG, CFL>(a, i);
        if constexpr (CFL) cfl_tile<MULTI>(a, tile, i, dti, sh);
    }
}

template <int W, int MAXDEG>
__device__ __forceinline__ void pass_b_cell_mixed(const PassArgs& a, int i, double dt, int inst) {
    double ud[NV];
    load_cell<W, NV>(a.U, i, ud);
    float u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = (float)ud[v];
    const float iv = a.iv32[i];
    float gi[2 * NV];
    load_cell<W, 2 * NV>(a.G32, i, gi);
    Kahan R[NV];
    int gnb[MAXDEG];
    double gn0[MAXDEG], gn1[MAXDEG];
    const float gma = (float)a.gamma;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gnb[k] = -1;
        gn0[k] = 0.0;
        gn1[k] = 0.0;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec32 fr = a.faces32[s.f & (SLOT_SIDE_R - 1)];
        const int j = s.nb;
        const bool ghost = j >= a.Nint;
        double und[NV];
        load_cell<W, NV>(a.U, j, und);
        float gj[2 * NV];
        if (ghost) {
#pragma unroll
            for (int q = 0; q < 2 * NV; q++) gj[q] = -gi[q];
            gnb[k] = j;
            // FP64 normal for the ghost state, from the FP64 record
            const FaceRec64 f64 = a.faces[s.f & (SLOT_SIDE_R - 1)];
            gn0[k] = f64.n0;
            gn1[k] = f64.n1;
        } else {
            load_cell<W, 2 * NV>(a.G32, j, gj);
        }
        const float ro0 = isR ? fr.rR0 : fr.rL0, ro1 = isR ? fr.rR1 : fr.rL1;
        const float rn0 = isR ? fr.rL0 : fr.rR0, rn1 = isR ? fr.rL1 : fr.rR1;
        float uo[NV], unb[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
            uo[v] = u[v] + (gi[2 * v] * ro0 + gi[2 * v + 1] * ro1);
            unb[v] = (float)und[v] + (gj[2 * v] * rn0 + gj[2 * v + 1] * rn1);
        }
        float F[NV];
        if (isR) rusanov(gma, unb, uo, fr.n0, fr.n1, F);
        else rusanov(gma, uo, unb, fr.n0, fr.n1, F);
#pragma unroll
        for (int v = 0; v < NV; v++) R[v].add(isR ? fr.A * F[v] : -(fr.A * F[v]));
    }
    double unew[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) unew[v] = ud[v] + dt * (R[v].value() * (double)iv);
    store_cell<W, NV>(a.Unext, i, unew);
    if (nonphysical(a.gamma, unew)) a.fault[inst] = 1;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (gnb[k] >= 0) {
            const BCEntry e = a.bctab[a.ghost_bc[gnb[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, gn0[k], gn1[k], ug);
            store_cell<W, NV>(a.Unext, gnb[k], ug);
        }
    }
}

template <int W, int MAXDEG, bool MULTI>
__global__ __launch_bounds__(256) void k_pass_b_mixed(PassArgs a) {
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + threadIdx.x;
        if (i >= a.Nint) continue;
        const int inst = MULTI ? instance_of(a.inst_off, a.ninst, i) : 0;
        pass_b_cell_mixed<W, MAXDEG>(a, i, a.dt[inst], inst);
    }
}

}  // namespace hcfd
