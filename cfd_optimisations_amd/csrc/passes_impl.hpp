// passes_impl.hpp -- kernel bodies of the two fused passes (included by
// passes_fp64.hip and passes_mixed.hip).
//
// One explicit step of Euler::run (src/Base/Euler.cpp:85-97) needs every cell's
// gradient and limiter before any face flux, so a step is two passes over the
// interior cells, each cell owning its own writes (no atomics, deterministic):
//
//   pass A  per cell: Green-Gauss gradient (GaussGreen.cpp:14-40), min/max
//           limiter (Limiter.cpp:19-77) [+ CFL bound]; then the MUSCL state of
//           every face of the cell on the cell's own side, U + phi (G . r)
//           (Flux.cpp:83-87), into cell-major storage FS[cell][slot] (each lane
//           writes one contiguous run); a cell next to a boundary also writes
//           the ghost's single slot, U_g + phi (-G . rR), since the ghost's
//           gradient is -G_L and its limiter phi_L (BoundaryConditions/Base.cpp:29-53).
//   pass B  per cell: Rusanov flux of each of its faces (Flux.cpp:52-109) from
//           its own FS[cell][k] and the neighbour's FS[nb][twin] (twin = the
//           face's slot in the neighbour's list), residual * 1/V, U += dt R
//           (Euler.cpp:107-113), and the ghost states of its boundary faces
//           from the new state (BoundaryConditions/*.cpp).
//
// Each cell sums its faces in ascending original face id -- the order the
// reference adds them -- and a face's flux is computed identically by both of
// its cells from the same stored face states, so in FP64 every value equals the
// reference's bit for bit (with -ffp-contract=off).
#pragma once
#include "device_math.hpp"
#include "passes.hpp"

// Minimum waves per SIMD the compiler must allow for each pass (bounds the
// VGPRs per lane): 512 VGPRs / waves.
#ifndef HCFD_WAVES_A
#define HCFD_WAVES_A 4
#endif
#ifndef HCFD_WAVES_B
#define HCFD_WAVES_B 4
#endif
#define HCFD_WAVES_A_ATTR __attribute__((amdgpu_waves_per_eu(HCFD_WAVES_A)))
#define HCFD_WAVES_B_ATTR __attribute__((amdgpu_waves_per_eu(HCFD_WAVES_B)))

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

__device__ __forceinline__ void store_fs(double* p, const double v[NV]) {
    double2* q = reinterpret_cast<double2*>(p);
    q[0] = make_double2(v[0], v[1]);
    q[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store_fs(float* p, const float v[NV]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

// ---------------------------------------------------------------------------
// Pass A, templated on the arithmetic type T of gradient/limiter/face states
// (double: the FP64 path; float: the mixed path, whose CFL bound stays FP64).
// ---------------------------------------------------------------------------
template <typename T>
struct PassATypes;
template <>
struct PassATypes<double> {
    using Rec = FaceRec64;
    __device__ static const Rec* rec(const PassArgs& a) { return a.faces; }
    __device__ static double* fs(const PassArgs& a) { return a.FS; }
    __device__ static double iv(const PassArgs& a, int i) { return a.iv[i]; }
};
template <>
struct PassATypes<float> {
    using Rec = FaceRec32;
    __device__ static const Rec* rec(const PassArgs& a) { return a.faces32; }
    __device__ static float* fs(const PassArgs& a) { return a.FS32; }
    __device__ static float iv(const PassArgs& a, int i) { return a.iv32[i]; }
};

template <typename T, int W, int MAXDEG, bool CFL>
__device__ __forceinline__ double pass_a_cell(const PassArgs& a, int i) {
    using PT = PassATypes<T>;
    double ud[NV];
    load_cell<W, NV>(a.U, i, ud);
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = (T)ud[v];
    const T iv = PT::iv(a, i);
    T g[NV][2], umin[NV], umax[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] = T(0);
        g[v][1] = T(0);
        umin[v] = u[v];
        umax[v] = u[v];
    }
    // CFL bound in FP64: lam = sum_f A_f (|u.n_f| + c)
    double ux = 0, uy = 0, c = 0, lam = 0.0;
    if constexpr (CFL) {
        const double gmo = a.gamma - 1;
        const double ir = 1.0 / ud[0];
        ux = ud[1] * ir;
        uy = ud[2] * ir;
        const double p = gmo * (ud[3] - 0.5 * (ud[1] * ux + ud[2] * uy));
        c = sqrt(a.gamma * p * ir);
    }
    T r0[MAXDEG], r1[MAXDEG];
    unsigned ghosts = 0;
    int deg = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        r0[k] = T(0);
        r1[k] = T(0);
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        deg = k + 1;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const int f = s.f & (SLOT_SIDE_R - 1);
        const typename PT::Rec fr = PT::rec(a)[f];
        const int nb = slot_nb(s);
        double und[NV];
        load_cell<W, NV>(a.U, nb, und);
        const T If = fr.If, omI = T(1) - If;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            // uf = U_L If + U_R (1 - If); G_L += uf A n, G_R -= uf A n
            const T un = (T)und[v];
            const T uf = isR ? (un * If + u[v] * omI) : (u[v] * If + un * omI);
            const T au = uf * fr.A;
            const T x0 = au * fr.n0, x1 = au * fr.n1;
            if (isR) {
                g[v][0] = g[v][0] - x0;
                g[v][1] = g[v][1] - x1;
            } else {
                g[v][0] = g[v][0] + x0;
                g[v][1] = g[v][1] + x1;
            }
            umin[v] = dmin(umin[v], un);
            umax[v] = dmax(umax[v], un);
        }
        r0[k] = isR ? fr.rR0 : fr.rL0;
        r1[k] = isR ? fr.rR1 : fr.rL1;
        if (nb >= a.Nint) ghosts |= 1u << k;
        if constexpr (CFL) lam += (double)fr.A * (fabs(ux * (double)fr.n0 + uy * (double)fr.n1) + c);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] *= iv;
        g[v][1] *= iv;
    }
    T phi[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        phi[v] = T(1);
        const T dmn = umin[v] - u[v], dmx = umax[v] - u[v];
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            if (k < deg) {
                const T d = g[v][0] * r0[k] + g[v][1] * r1[k];
                phi[v] = dmin(phi[v], limiter_face(d, dmn, dmx));
            }
        }
    }
    // MUSCL face states on this cell's side, U + phi (G . r)
    T* FS = PT::fs(a) + (long long)i * (MAXDEG * NV);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (k < deg) {
            T uf[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) uf[v] = u[v] + phi[v] * (g[v][0] * r0[k] + g[v][1] * r1[k]);
            store_fs(FS + k * NV, uf);
            if ((ghosts >> k) & 1u) {
                // the ghost side (R) of a boundary face: U_g + phi_L ((-G_L) . rR)
                const Slot s = a.slots[(long long)k * a.NcPad + i];
                const typename PT::Rec fr = PT::rec(a)[s.f & (SLOT_SIDE_R - 1)];
                const int gc = slot_nb(s);
                double ugd[NV];
                load_cell<W, NV>(a.U, gc, ugd);
                T ufg[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) ufg[v] = (T)ugd[v] + phi[v] * ((-g[v][0]) * fr.rR0 + (-g[v][1]) * fr.rR1);
                store_fs(PT::fs(a) + (long long)gc * (MAXDEG * NV), ufg);
            }
        }
    }
    if constexpr (CFL) return 1.0 / ((double)iv * lam);
    return 0.0;
}

// Pass A with four lanes per cell, lane v handling conserved variable v
// (gradient, limiter and face states are independent per variable; the CFL
// bound needs the whole state, exchanged within the quad).  Same arithmetic per
// variable as pass_a_cell, so the results are identical.
template <typename T, int W, int MAXDEG, bool CFL>
__device__ __forceinline__ double pass_a_cell_quad(const PassArgs& a, int i, int v) {
    using PT = PassATypes<T>;
    const double udv = a.U[cidx<W, NV>(i, v)];
    const T u = (T)udv;
    const T iv = PT::iv(a, i);
    T g0 = T(0), g1 = T(0), umin = u, umax = u;
    double ux = 0, uy = 0, c = 0, lam = 0.0;
    if constexpr (CFL) {
        const int q = threadIdx.x & ~3;
        const double u0 = __shfl(udv, q + 0, 64), u1 = __shfl(udv, q + 1, 64);
        const double u2 = __shfl(udv, q + 2, 64), u3 = __shfl(udv, q + 3, 64);
        const double gmo = a.gamma - 1;
        const double ir = 1.0 / u0;
        ux = u1 * ir;
        uy = u2 * ir;
        const double p = gmo * (u3 - 0.5 * (u1 * ux + u2 * uy));
        c = sqrt(a.gamma * p * ir);
    }
    T r0[MAXDEG], r1[MAXDEG];
    unsigned ghosts = 0;
    int deg = 0;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        r0[k] = T(0);
        r1[k] = T(0);
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        deg = k + 1;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const int f = s.f & (SLOT_SIDE_R - 1);
        const typename PT::Rec fr = PT::rec(a)[f];
        const int nb = slot_nb(s);
        const T un = (T)a.U[cidx<W, NV>(nb, v)];
        const T If = fr.If, omI = T(1) - If;
        const T uf = isR ? (un * If + u * omI) : (u * If + un * omI);
        const T au = uf * fr.A;
        const T x0 = au * fr.n0, x1 = au * fr.n1;
        if (isR) {
            g0 = g0 - x0;
            g1 = g1 - x1;
        } else {
            g0 = g0 + x0;
            g1 = g1 + x1;
        }
        umin = dmin(umin, un);
        umax = dmax(umax, un);
        r0[k] = isR ? fr.rR0 : fr.rL0;
        r1[k] = isR ? fr.rR1 : fr.rL1;
        if (nb >= a.Nint) ghosts |= 1u << k;
        if constexpr (CFL) lam += (double)fr.A * (fabs(ux * (double)fr.n0 + uy * (double)fr.n1) + c);
    }
    g0 *= iv;
    g1 *= iv;
    T phi = T(1);
    const T dmn = umin - u, dmx = umax - u;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++)
        if (k < deg) phi = dmin(phi, limiter_face(g0 * r0[k] + g1 * r1[k], dmn, dmx));
    T* FS = PT::fs(a);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (k < deg) {
            FS[((long long)i * MAXDEG + k) * NV + v] = u + phi * (g0 * r0[k] + g1 * r1[k]);
            if ((ghosts >> k) & 1u) {
                const Slot s = a.slots[(long long)k * a.NcPad + i];
                const typename PT::Rec fr = PT::rec(a)[s.f & (SLOT_SIDE_R - 1)];
                const int gc = slot_nb(s);
                const T ug = (T)a.U[cidx<W, NV>(gc, v)];
                FS[(long long)gc * (MAXDEG * NV) + v] = ug + phi * ((-g0) * fr.rR0 + (-g1) * fr.rR1);
            }
        }
    }
    if constexpr (CFL) return v == 0 ? 1.0 / ((double)iv * lam) : INFINITY;
    return 0.0;
}

// LPC lanes per cell (1: one cell per lane; 4: one variable per lane).
template <typename T, int W, int MAXDEG, bool CFL, bool MULTI, int LPC>
__global__ __launch_bounds__(256) HCFD_WAVES_A_ATTR void k_pass_a(PassArgs a) {
    __shared__ double sh[256 / WAVE];
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + (int)threadIdx.x / LPC;
        double dti = INFINITY;
        if (i < a.Nint) {
            if constexpr (LPC == 4) dti = pass_a_cell_quad<T, W, MAXDEG, CFL>(a, i, threadIdx.x & 3);
            else dti = pass_a_cell<T, W, MAXDEG, CFL>(a, i);
        }
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
// Pass B, FP64
// ---------------------------------------------------------------------------
template <int W, int MAXDEG>
__device__ __forceinline__ void pass_b_cell_fp64(const PassArgs& a, int i, double dt, int inst) {
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const double iv = a.iv[i];
    double R[NV] = {0.0, 0.0, 0.0, 0.0};
    int gg[MAXDEG];
    double gn0[MAXDEG], gn1[MAXDEG];
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gg[k] = -1;
        gn0[k] = 0.0;
        gn1[k] = 0.0;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const int f = s.f & (SLOT_SIDE_R - 1);
        const double2* qo = reinterpret_cast<const double2*>(a.FS + ((long long)i * MAXDEG + k) * NV);
        const double2* qn =
            reinterpret_cast<const double2*>(a.FS + ((long long)slot_nb(s) * MAXDEG + slot_twin(s)) * NV);
        const double2 o0 = qo[0], o1 = qo[1], n0 = qn[0], n1 = qn[1];
        const double2 l0 = isR ? n0 : o0, l1 = isR ? n1 : o1, r0 = isR ? o0 : n0, r1 = isR ? o1 : n1;
        const double UL[NV] = {l0.x, l0.y, l1.x, l1.y};
        const double UR[NV] = {r0.x, r0.y, r1.x, r1.y};
        const FaceN64 fn = a.fn[f];
        double F[NV];
        rusanov(a.gamma, UL, UR, fn.n0, fn.n1, F);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (isR) R[v] = R[v] + fn.A * F[v];
            else R[v] = R[v] - fn.A * F[v];
        }
        if (fn.ghost >= 0) {
            gg[k] = fn.ghost;
            gn0[k] = fn.n0;
            gn1[k] = fn.n1;
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
        if (gg[k] >= 0) {
            const BCEntry e = a.bctab[a.ghost_bc[gg[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, gn0[k], gn1[k], ug);
            store_cell<W, NV>(a.Unext, gg[k], ug);
        }
    }
}

// ---------------------------------------------------------------------------
// Pass B, FP32-mixed: flux in FP32 from FP32 face states, residual summed with
// Kahan compensation, update and ghost states in FP64.
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

template <int W, int MAXDEG>
__device__ __forceinline__ void pass_b_cell_mixed(const PassArgs& a, int i, double dt, int inst) {
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const float iv = a.iv32[i];
    const float gma = (float)a.gamma;
    Kahan R[NV];
    int gg[MAXDEG], gf[MAXDEG];
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gg[k] = -1;
        gf[k] = 0;
        const Slot s = a.slots[(long long)k * a.NcPad + i];
        if (s.f < 0) continue;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const int f = s.f & (SLOT_SIDE_R - 1);
        const float4 qo = *reinterpret_cast<const float4*>(a.FS32 + ((long long)i * MAXDEG + k) * NV);
        const float4 qn =
            *reinterpret_cast<const float4*>(a.FS32 + ((long long)slot_nb(s) * MAXDEG + slot_twin(s)) * NV);
        const float4 qL = isR ? qn : qo, qR = isR ? qo : qn;
        const float UL[NV] = {qL.x, qL.y, qL.z, qL.w};
        const float UR[NV] = {qR.x, qR.y, qR.z, qR.w};
        const FaceN32 fn = a.fn32[f];
        float F[NV];
        rusanov(gma, UL, UR, fn.n0, fn.n1, F);
#pragma unroll
        for (int v = 0; v < NV; v++) R[v].add(isR ? fn.A * F[v] : -(fn.A * F[v]));
        if (fn.ghost >= 0) {
            gg[k] = fn.ghost;
            gf[k] = f;
        }
    }
    double unew[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) unew[v] = u[v] + dt * (R[v].value() * (double)iv);
    store_cell<W, NV>(a.Unext, i, unew);
    if (nonphysical(a.gamma, unew)) a.fault[inst] = 1;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (gg[k] >= 0) {
            const FaceN64 f64 = a.fn[gf[k]];  // FP64 normal for the ghost state
            const BCEntry e = a.bctab[a.ghost_bc[gg[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, f64.n0, f64.n1, ug);
            store_cell<W, NV>(a.Unext, gg[k], ug);
        }
    }
}

template <bool MIXED, int W, int MAXDEG, bool MULTI>
__global__ __launch_bounds__(256) HCFD_WAVES_B_ATTR void k_pass_b(PassArgs a) {
    int t0, step, tend;
    tile_walk(a, t0, step, tend);
    for (int tile = t0; tile < tend; tile += step) {
        const int i = tile * a.tile_cells + threadIdx.x;
        if (i >= a.Nint) continue;
        const int inst = MULTI ? instance_of(a.inst_off, a.ninst, i) : 0;
        if constexpr (MIXED) pass_b_cell_mixed<W, MAXDEG>(a, i, a.dt[inst], inst);
        else pass_b_cell_fp64<W, MAXDEG>(a, i, a.dt[inst], inst);
    }
}

}  // namespace hcfd
