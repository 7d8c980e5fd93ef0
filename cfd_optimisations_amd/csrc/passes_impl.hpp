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
//           (Flux.cpp:83-87), into slot-major planes FS[slot][cell]; a cell next
//           to a boundary also writes the ghost's single slot, U_g + phi (-G . rR),
//           since the ghost's gradient is -G_L and its limiter phi_L
//           (BoundaryConditions/Base.cpp:29-53).
//   pass B  per cell: Rusanov flux of each of its faces (Flux.cpp:52-109) from
//           its own FS[k][cell] and the neighbour's FS[twin][nb] (twin = the
//           face's slot in the neighbour's list), residual * 1/V, U += dt R
//           (Euler.cpp:107-113) [+ the SSP-RK stage combination with the step's
//           starting state], and the ghost states of its boundary faces from the
//           new state (BoundaryConditions/{Symmetry,Wall,Inlet,Outlet}.cpp).
//
// Geometry is read from per-slot planes (cell-major, so each load of a wave is
// one contiguous run) with the normal oriented out of the cell.  Each cell sums
// its faces in ascending original face id -- the order the reference adds
// them -- and the arithmetic stays the reference's value for value:
//   - gradient: G_L += (uf A) n and G_R -= (uf A) n equal G += (uf A) n_out,
//     since x * (-y) = -(x y) and g - x = g + (-x) exactly; uf's two products
//     are the same pair in either order (addition commutes exactly);
//   - flux: Rusanov is exactly antisymmetric, F(U_R, U_L, -n) = -F(U_L, U_R, n)
//     term by term (negation and the commutativity of +, * and max are exact),
//     so every cell evaluates the flux in its outward frame from its own state
//     and its neighbour's, and R -= A F_out equals the reference's R_L -= A F /
//     R_R += A F;
//   - the ghost side of a boundary face: rR = -rL there (Domain.cpp:84-88), so
//     (-G) . rR = G . rL exactly.
// With -ffp-contract=off the FP64 results equal the reference's bit for bit.
#pragma once
#include <type_traits>

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

// CFL partials (maximum rate, see cfl_rate) of one tile of cells [c0, c1]: the
// tile's first and last instance get block maxima; cells of any instance
// strictly inside the tile use atomics.  Lanes without a cell pass i >= Nint.
template <bool MULTI>
__device__ __forceinline__ void cfl_tile_range(const PassArgs& a, int tile, int c0, int c1, int i, double rate,
                                               double* sh) {
    if constexpr (!MULTI) {
        const double m = block_max(rate, sh);
        if (threadIdx.x == 0) a.pmin_first[tile] = m;
    } else {
        const int bf = instance_from(a.inst_off, a.chunk_inst, c0);
        const int bl = instance_from(a.inst_off, a.chunk_inst, c1);
        double vf = 0.0, vl = 0.0;
        if (i < a.Nint) {
            const int bi = instance_from(a.inst_off, a.chunk_inst, i);
            if (bi == bf) vf = rate;
            else if (bi == bl) vl = rate;
            else atomicMax(&a.dtacc[bi], dbits(rate));
        }
        const double mf = block_max(vf, sh);
        const double ml = block_max(vl, sh);
        if (threadIdx.x == 0) {
            a.pmin_first[tile] = mf;
            a.pmin_last[tile] = ml;
        }
    }
}

// Tiles of tile_cells consecutive cells (the two-pass kernels).
template <bool MULTI>
__device__ __forceinline__ void cfl_tile(const PassArgs& a, int tile, int i, double rate, double* sh) {
    const int c0 = tile * a.tile_cells;
    cfl_tile_range<MULTI>(a, tile, c0, min(c0 + a.tile_cells, a.Nint) - 1, i, rate, sh);
}

// Face-state planes: FP64 slot k of cell i is two double2 at planes 2k, 2k+1;
// FP32 slot k is one float4 at plane k (plane pitch NcPad).
__device__ __forceinline__ void store_fs(double* fs, long long P, int k, int i, const double v[NV]) {
    double2* q = reinterpret_cast<double2*>(fs);
    q[(2 * k) * P + i] = make_double2(v[0], v[1]);
    q[(2 * k + 1) * P + i] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store_fs(float* fs, long long P, int k, int i, const float v[NV]) {
    reinterpret_cast<float4*>(fs)[k * P + i] = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void load_fs(const double* fs, long long P, int k, int i, double v[NV]) {
    const double2* q = reinterpret_cast<const double2*>(fs);
    const double2 a = q[(2 * k) * P + i], b = q[(2 * k + 1) * P + i];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load_fs(const float* fs, long long P, int k, int i, float v[NV]) {
    const float4 a = reinterpret_cast<const float4*>(fs)[k * P + i];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
// Element index of component v of face state (k, i) in the planes above.
template <typename T>
__device__ __forceinline__ long long fs_index(long long P, int k, int i, int v) {
    if constexpr (sizeof(T) == 8) return (((2 * k + (v >> 1)) * P + i) << 1) + (v & 1);
    else return ((k * P + i) << 2) + v;
}
// Geometry plane p of slot k of cell i (MAXDEG slot planes per quantity, in
// 64-cell blocks: geo_index; the pitch argument is not needed by that layout).
template <int MAXDEG, typename T>
__device__ __forceinline__ T geo(const T* __restrict__ g, long long, int p, int k, int i) {
    return g[geo_index(MAXDEG, p, k, (unsigned)i)];
}

// Slot k of cell i in the neighbour table.
__device__ __forceinline__ int cnb_at(const PassArgs& a, int k, int i) {
    return plane_at(a.cnb + (long long)k * a.NcPad, i);
}

// ---------------------------------------------------------------------------
// Pass A, templated on the arithmetic type T of gradient/limiter/face states
// (double: the FP64 path; float: the mixed path, whose CFL bound stays FP64).
// ---------------------------------------------------------------------------
template <typename T>
struct PassATypes;
template <>
struct PassATypes<double> {
    __device__ static const double* geo(const PassArgs& a) { return a.cgeo; }
    __device__ static double* fs(const PassArgs& a) { return a.FS; }
    __device__ static double iv(const PassArgs& a, int i) { return plane_at(a.iv, i); }
    __device__ static const double* ivp(const PassArgs& a) { return a.iv; }
    __device__ static const double* ring_geo(const PassArgs& a) { return a.hgeo; }
    __device__ static const double* ring_iv(const PassArgs& a) { return a.hiv; }
};
template <>
struct PassATypes<float> {
    __device__ static const float* geo(const PassArgs& a) { return a.cgeo32; }
    __device__ static float* fs(const PassArgs& a) { return a.FS32; }
    __device__ static float iv(const PassArgs& a, int i) { return plane_at(a.iv32, i); }
    __device__ static const float* ivp(const PassArgs& a) { return a.iv32; }
    __device__ static const float* ring_geo(const PassArgs& a) { return a.hgeo32; }
    __device__ static const float* ring_iv(const PassArgs& a) { return a.hiv32; }
};

// Tile staging: the block's cells [tb, tb + tc) have their state in LDS
// (two double2 per cell); a neighbour inside the tile is read from there.
__device__ __forceinline__ void lds_cell(const double2* __restrict__ s, int l, double out[NV]) {
    const double2 x = s[2 * l], y = s[2 * l + 1];
    out[0] = x.x; out[1] = x.y; out[2] = y.x; out[3] = y.y;
}
template <int W>
__device__ __forceinline__ void tile_or_global_cell(const PassArgs& a, const double2* __restrict__ s, int tb, int tc,
                                                    int j, double out[NV]) {
    const unsigned l = (unsigned)(j - tb);
    if (l < (unsigned)tc) lds_cell(s, (int)l, out);
    else load_cell<W, NV>(a.U, j, out);
}

template <typename T, int W, int MAXDEG, bool CFL>
__device__ __forceinline__ double pass_a_cell(const PassArgs& a, int i, const double2* __restrict__ sU, int tb,
                                              int tc) {
    using PT = PassATypes<T>;
    const long long P = a.NcPad;
    const T* __restrict__ gp = PT::geo(a);
    double ud[NV];
    lds_cell(sU, i - tb, ud);
    // FP64: the state itself; FP32-mixed: differences from it (u = 0, see
    // rusanov_delta), so the state is never rounded to FP32
    constexpr bool DELTA = sizeof(T) == 4;
    T u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = DELTA ? T(0) : (T)ud[v];
    const T iv = PT::iv(a, i);
    const int side = plane_at(a.cside, i);
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
        const int nbx = cnb_at(a, k, i);
        if (nbx == SLOT_EMPTY) continue;
        deg = k + 1;
        const int nb = nbx & SLOT_NB_MASK;
        const T A = geo<MAXDEG>(gp, P, GEO_A, k, i);
        const T n0 = geo<MAXDEG>(gp, P, GEO_N0, k, i);
        const T n1 = geo<MAXDEG>(gp, P, GEO_N1, k, i);
        const T If = geo<MAXDEG>(gp, P, GEO_IF, k, i);
        r0[k] = geo<MAXDEG>(gp, P, GEO_R0, k, i);
        r1[k] = geo<MAXDEG>(gp, P, GEO_R1, k, i);
        double und[NV];
        tile_or_global_cell<W>(a, sU, tb, tc, nb, und);
        // uf = U_L If + U_R (1 - If): w_own multiplies this cell's state
        const T omI = T(1) - If;
        const bool isR = (side >> k) & 1;
        const T wo = isR ? omI : If, wn = isR ? If : omI;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            // mixed: un = u_nb - u and uf - u = (u_nb - u) wn
            const T un = DELTA ? (T)(und[v] - ud[v]) : (T)und[v];
            const T uf = DELTA ? un * wn : madd(u[v], wo, un * wn);
            // mixed: n0, n1 hold A n
            const T au = DELTA ? uf : uf * A;
            g[v][0] = madd(au, n0, g[v][0]);
            g[v][1] = madd(au, n1, g[v][1]);
            umin[v] = dmin(umin[v], un);
            umax[v] = dmax(umax[v], un);
        }
        if (nb >= a.Nint) ghosts |= 1u << k;
        if constexpr (CFL) lam += cfl_face(ux, uy, c, A, n0, n1);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        g[v][0] *= iv;
        g[v][1] *= iv;
    }
    T phi[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        T dpos = T(0), dneg = T(0);
#pragma unroll
        for (int k = 0; k < MAXDEG; k++) {
            if (k < deg) {
                const T d = dot_r(g[v][0], g[v][1], r0[k], r1[k]);
                dpos = dmax(dpos, d);
                dneg = dmin(dneg, d);
            }
        }
        phi[v] = limiter_phi(dpos, dneg, umin[v] - u[v], umax[v] - u[v]);
    }
    // MUSCL face states on this cell's side, U + phi (G . r)
    T* FS = PT::fs(a);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (k < deg) {
            T d[NV], uf[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                d[v] = dot_r(g[v][0], g[v][1], r0[k], r1[k]);
                uf[v] = madd(phi[v], d[v], u[v]);
            }
            store_fs(FS, P, k, i, uf);
            // mixed: the ghost's increment is this cell's (pass B uses it directly)
            if (!DELTA && ((ghosts >> k) & 1u)) {
                // the ghost side (R) of a boundary face: U_g + phi_L ((-G_L) . rR), rR = -rL
                const int gc = cnb_at(a, k, i) & SLOT_NB_MASK;
                double ugd[NV];
                tile_or_global_cell<W>(a, sU, tb, tc, gc, ugd);
                T ufg[NV];
#pragma unroll
                for (int v = 0; v < NV; v++) ufg[v] = madd(phi[v], d[v], (T)ugd[v]);
                store_fs(FS, P, 0, gc, ufg);
            }
        }
    }
    if constexpr (CFL) return cfl_rate((double)iv * lam);
    return 0.0;
}

// Pass A with four lanes per cell, lane v handling conserved variable v
// (gradient, limiter and face states are independent per variable; the CFL
// bound needs the whole state, exchanged within the quad).  Same arithmetic per
// variable as pass_a_cell, so the results are identical.
template <typename T, int W, int MAXDEG, bool CFL>
__device__ __forceinline__ double pass_a_cell_quad(const PassArgs& a, int i, int v) {
    using PT = PassATypes<T>;
    const long long P = a.NcPad;
    const T* __restrict__ gp = PT::geo(a);
    const double udv = a.U[cidx<W, NV>(i, v)];
    constexpr bool DELTA = sizeof(T) == 4;  // mixed: differences from the state (see pass_a_cell)
    const T u = DELTA ? T(0) : (T)udv;
    const T iv = PT::iv(a, i);
    const int side = plane_at(a.cside, i);
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
        const int nbx = cnb_at(a, k, i);
        if (nbx == SLOT_EMPTY) continue;
        deg = k + 1;
        const int nb = nbx & SLOT_NB_MASK;
        const T A = geo<MAXDEG>(gp, P, GEO_A, k, i);
        const T n0 = geo<MAXDEG>(gp, P, GEO_N0, k, i);
        const T n1 = geo<MAXDEG>(gp, P, GEO_N1, k, i);
        const T If = geo<MAXDEG>(gp, P, GEO_IF, k, i);
        r0[k] = geo<MAXDEG>(gp, P, GEO_R0, k, i);
        r1[k] = geo<MAXDEG>(gp, P, GEO_R1, k, i);
        const double unb = a.U[cidx<W, NV>(nb, v)];
        const T un = DELTA ? (T)(unb - udv) : (T)unb;
        const T omI = T(1) - If;
        const bool isR = (side >> k) & 1;
        const T uf = DELTA ? un * (isR ? If : omI) : madd(u, isR ? omI : If, un * (isR ? If : omI));
        const T au = DELTA ? uf : uf * A;  // mixed: n0, n1 hold A n
        g0 = madd(au, n0, g0);
        g1 = madd(au, n1, g1);
        umin = dmin(umin, un);
        umax = dmax(umax, un);
        if (nb >= a.Nint) ghosts |= 1u << k;
        if constexpr (CFL) lam += cfl_face(ux, uy, c, A, n0, n1);
    }
    g0 *= iv;
    g1 *= iv;
    T dpos = T(0), dneg = T(0);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (k < deg) {
            const T d = dot_r(g0, g1, r0[k], r1[k]);
            dpos = dmax(dpos, d);
            dneg = dmin(dneg, d);
        }
    }
    const T phi = limiter_phi(dpos, dneg, umin - u, umax - u);
    T* FS = PT::fs(a);
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (k < deg) {
            const T d = dot_r(g0, g1, r0[k], r1[k]);
            FS[fs_index<T>(P, k, i, v)] = madd(phi, d, u);
            if (!DELTA && ((ghosts >> k) & 1u)) {
                const int gc = cnb_at(a, k, i) & SLOT_NB_MASK;
                const T ug = (T)a.U[cidx<W, NV>(gc, v)];
                FS[fs_index<T>(P, 0, gc, v)] = madd(phi, d, ug);
            }
        }
    }
    if constexpr (CFL) return v == 0 ? cfl_rate((double)iv * lam) : 0.0;
    return 0.0;
}

// LPC lanes per cell (1: one cell per lane; 4: one variable per lane).
// With one lane per cell the tile's states are staged in LDS first.
template <typename T, int W, int MAXDEG, bool CFL, bool MULTI, int LPC>
__global__ __launch_bounds__(256) HCFD_WAVES_A_ATTR void k_pass_a(PassArgs a) {
    __shared__ double sh[256 / WAVE];
    __shared__ double2 sU[LPC == 1 ? 2 * 256 : 1];
    int tfirst, step, tend;
    tile_walk(a, tfirst, step, tend);
    for (int tile = tfirst; tile < tend; tile += step) {
        const int tb = tile * a.tile_cells;
        const int i = tb + (int)threadIdx.x / LPC;
        if constexpr (LPC == 1) {
            __syncthreads();  // the previous tile's readers are done
            if (i < a.Nc) {
                double u[NV];
                load_cell<W, NV>(a.U, i, u);
                sU[2 * threadIdx.x] = make_double2(u[0], u[1]);
                sU[2 * threadIdx.x + 1] = make_double2(u[2], u[3]);
            }
            __syncthreads();
        }
        double rate = 0.0;
        if (i < a.Nint) {
            if constexpr (LPC == 4) rate = pass_a_cell_quad<T, W, MAXDEG, CFL>(a, i, threadIdx.x & 3);
            else rate = pass_a_cell<T, W, MAXDEG, CFL>(a, i, sU, tb, min(a.tile_cells, a.Nc - tb));
        }
        if constexpr (CFL) cfl_tile<MULTI>(a, tile, i, rate, sh);
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

__device__ __forceinline__ bool nonphysical(double, const double u[NV]) {
    // rho > 0 and p = (gamma - 1)(E - |m|^2 / (2 rho)) > 0, i.e. 2 rho E > |m|^2
    // (multiplied through by 2 rho > 0: no division); NaN fails either test
    return !(u[0] > 0.0) || !(2.0 * u[0] * u[3] > u[1] * u[1] + u[2] * u[2]);
}

// ---------------------------------------------------------------------------
// Pass B, FP64
// ---------------------------------------------------------------------------
// Face states of the block's tile staged in LDS: FP64 slot k of tile cell l is
// s[(2k) * 256 + l], s[(2k + 1) * 256 + l]; FP32 slot k is s[k * 256 + l].
constexpr int LDS_TILE = 256;
__device__ __forceinline__ void lds_fs(const double2* __restrict__ s, int k, int l, double v[NV]) {
    const double2 x = s[(2 * k) * LDS_TILE + l], y = s[(2 * k + 1) * LDS_TILE + l];
    v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
}
__device__ __forceinline__ void lds_fs(const float4* __restrict__ s, int k, int l, float v[NV]) {
    const float4 x = s[k * LDS_TILE + l];
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
}
template <typename T, typename S>
__device__ __forceinline__ void tile_or_global_fs(const T* fs, long long P, const S* __restrict__ s, int tb, int tc,
                                                  int k, int j, T v[NV]) {
    const unsigned l = (unsigned)(j - tb);
    if (l < (unsigned)tc) lds_fs(s, k, (int)l, v);
    else load_fs(fs, P, k, j, v);
}

// The step's clocks, on the device in both dt modes (the same FP64 additions
// as Euler::run's t += dt, src/Base/Euler.cpp:96): block 0 of the first stage.
__device__ __forceinline__ void advance_clocks(const PassArgs& a) {
    if (a.advance && blockIdx.x == 0)
        for (int b = threadIdx.x; b < a.ninst; b += blockDim.x) {
            a.tacc[b] += a.dt[b];
            a.lastdt[b] = a.dt[b];
        }
}

// Stage combination of a multi-stage step: unew = rk_a * U0 + rk_b * unew, the
// two products rounded separately (no contraction), as the Shu-Osher form reads.
template <int W>
__device__ __forceinline__ void rk_combine(const PassArgs& a, int i, double unew[NV]) {
    double u0[NV];
    load_cell<W, NV>(a.U0, i, u0);
#pragma unroll
    for (int v = 0; v < NV; v++) unew[v] = a.rk_a * u0[v] + a.rk_b * unew[v];
}

template <int W, int MAXDEG>
__device__ __forceinline__ void pass_b_cell_fp64(const PassArgs& a, int i, double dt, int inst,
                                                 const double2* __restrict__ sF, int tb, int tc) {
    const long long P = a.NcPad;
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const double iv = a.iv[i];
    double R[NV] = {0.0, 0.0, 0.0, 0.0};
    int gg[MAXDEG];
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gg[k] = -1;
        const int nbx = cnb_at(a, k, i);
        if (nbx == SLOT_EMPTY) continue;
        const int nb = nbx & SLOT_NB_MASK;
        const int tw = (int)((unsigned)nbx >> SLOT_TWIN_SHIFT);
        double Uo[NV], Un[NV];
        lds_fs(sF, k, i - tb, Uo);
        tile_or_global_fs(a.FS, P, sF, tb, tc, tw, nb, Un);
        const double A = geo<MAXDEG>(a.cgeo, P, GEO_A, k, i);
        const double n0 = geo<MAXDEG>(a.cgeo, P, GEO_N0, k, i);
        const double n1 = geo<MAXDEG>(a.cgeo, P, GEO_N1, k, i);
        double F[NV];
        rusanov(a.gamma, Uo, Un, n0, n1, F);  // outward flux of this cell
#pragma unroll
        for (int v = 0; v < NV; v++) R[v] = R[v] - A * F[v];
        if (nb >= a.Nint) gg[k] = nb;
    }
    double unew[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        R[v] = R[v] * iv;
        unew[v] = u[v] + dt * R[v];
    }
    if (a.U0) rk_combine<W>(a, i, unew);
    store_cell<W, NV>(a.Unext, i, unew);
    if (nonphysical(a.gamma, unew)) a.fault[inst] = 1;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (gg[k] >= 0) {
            // boundary faces have the interior cell on the L side: n_out = n
            const BCEntry e = a.bctab[a.ghost_bc[gg[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, geo<MAXDEG>(a.cgeo, P, GEO_N0, k, i), geo<MAXDEG>(a.cgeo, P, GEO_N1, k, i),
                        ug);
            store_cell<W, NV>(a.Unext, gg[k], ug);
        }
    }
}

// ---------------------------------------------------------------------------
// Pass B, FP32-mixed: flux in FP32 in difference form (rusanov_delta) from the
// FP32 face-state increments and the FP64 neighbour differences, residual
// summed in FP64 (each FP32 term exactly, one rounding per addition), update
// and ghost states in FP64.
// ---------------------------------------------------------------------------
template <int W, int MAXDEG>
__device__ __forceinline__ void pass_b_cell_mixed(const PassArgs& a, int i, double dt, int inst,
                                                  const float4* __restrict__ sF, int tb, int tc) {
    const long long P = a.NcPad;
    double u[NV];
    load_cell<W, NV>(a.U, i, u);
    const float iv = a.iv32[i];
    const float gma = (float)a.gamma;
    const FluxBase fb = flux_base(gma, u);
    double R[NV] = {0.0, 0.0, 0.0, 0.0};
    int gg[MAXDEG];
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        gg[k] = -1;
        const int nbx = cnb_at(a, k, i);
        if (nbx == SLOT_EMPTY) continue;
        const int nb = nbx & SLOT_NB_MASK;
        const int tw = (int)((unsigned)nbx >> SLOT_TWIN_SHIFT);
        // increments of the two face states: own phi (G . r); the neighbour's
        // u_nb - u plus its own increment (a ghost's increment is this cell's)
        float dO[NV], dN[NV];
        lds_fs(sF, k, i - tb, dO);
        double un[NV];
        load_cell<W, NV>(a.U, nb, un);
        if (nb >= a.Nint) {
#pragma unroll
            for (int v = 0; v < NV; v++) dN[v] = dO[v];
        } else {
            tile_or_global_fs(a.FS32, P, sF, tb, tc, tw, nb, dN);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) dN[v] += (float)(un[v] - u[v]);
        float F[NV];  // 2 A (F - E(u).n): the mixed geometry planes hold A n
        rusanov_delta(gma, fb, dO, dN, geo<MAXDEG>(a.cgeo32, P, GEO_N0, k, i), geo<MAXDEG>(a.cgeo32, P, GEO_N1, k, i),
                      F);
#pragma unroll
        for (int v = 0; v < NV; v++) R[v] = R[v] - (double)F[v];
        if (nb >= a.Nint) gg[k] = nb;
    }
    double unew[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) unew[v] = u[v] + dt * ((0.5 * R[v]) * (double)iv);
    if (a.U0) rk_combine<W>(a, i, unew);
    store_cell<W, NV>(a.Unext, i, unew);
    if (nonphysical(a.gamma, unew)) a.fault[inst] = 1;
#pragma unroll
    for (int k = 0; k < MAXDEG; k++) {
        if (gg[k] >= 0) {
            // FP64 normal for the ghost state (interior cell on the L side)
            const BCEntry e = a.bctab[a.ghost_bc[gg[k] - a.Nint]];
            double ug[NV];
            ghost_state(e, a.gamma, unew, geo<MAXDEG>(a.cgeo, P, GEO_N0, k, i), geo<MAXDEG>(a.cgeo, P, GEO_N1, k, i),
                        ug);
            store_cell<W, NV>(a.Unext, gg[k], ug);
        }
    }
}

// The tile's face states (every slot of every cell of the tile, ghosts
// included) are staged in LDS; neighbours inside the tile are read from there.
template <bool MIXED, int W, int MAXDEG, bool MULTI>
__global__ __launch_bounds__(256) HCFD_WAVES_B_ATTR void k_pass_b(PassArgs a) {
    using S = typename std::conditional<MIXED, float4, double2>::type;
    constexpr int PLANES = MIXED ? MAXDEG : 2 * MAXDEG;
    __shared__ S sF[PLANES * LDS_TILE];
    const long long P = a.NcPad;
    advance_clocks(a);
    int tfirst, step, tend;
    tile_walk(a, tfirst, step, tend);
    for (int tile = tfirst; tile < tend; tile += step) {
        const int tb = tile * a.tile_cells;
        const int i = tb + threadIdx.x;
        __syncthreads();  // the previous tile's readers are done
        if (i < a.Nc) {
            const S* src = reinterpret_cast<const S*>(MIXED ? (const void*)a.FS32 : (const void*)a.FS);
#pragma unroll
            for (int q = 0; q < PLANES; q++) sF[q * LDS_TILE + threadIdx.x] = src[q * P + i];
        }
        __syncthreads();
        if (i >= a.Nint) continue;
        const int tc = min(a.tile_cells, a.Nc - tb);
        const int inst = MULTI ? instance_from(a.inst_off, a.chunk_inst, i) : 0;
        if constexpr (MIXED) pass_b_cell_mixed<W, MAXDEG>(a, i, a.dt[inst], inst, sF, tb, tc);
        else pass_b_cell_fp64<W, MAXDEG>(a, i, a.dt[inst], inst, sF, tb, tc);
    }
}

}  // namespace hcfd
