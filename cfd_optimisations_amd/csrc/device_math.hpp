// device_math.hpp -- per-face arithmetic of the residual, for gfx950 kernels.
//
// The FP64 functions keep the reference's expression order term by term so
// that, compiled with -ffp-contract=off, each value is bit-identical to
// Base's: EulerFlux (src/Base/Flux.cpp:89-97), Rusanov (:52-81) and the MUSCL
// face state (:83-87).  The FP32 variants serve the mixed path.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace hcfd {

__device__ __forceinline__ int slot_nb(const Slot& s) { return s.nb & SLOT_NB_MASK; }
__device__ __forceinline__ int slot_twin(const Slot& s) { return (int)((unsigned)s.nb >> SLOT_TWIN_SHIFT); }

template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return fmin(a, b); }
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return fmax(a, b); }
__device__ __forceinline__ float dmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }

// Reciprocal, quotient and square root: IEEE-correct in FP64 (the reference's
// values); in FP32 (the mixed path) the hardware instructions v_rcp_f32 and
// v_sqrt_f32 (1 ulp), a fifth of the instructions of the correctly rounded
// expansions and well inside the mixed path's error budget.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double quot(double a, double b) { return a / b; }
__device__ __forceinline__ float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double vsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float vsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// a * b + c.  FP64: product and sum rounded separately, the reference's
// arithmetic (IEEE addition commutes exactly, so x + a * b written as
// madd(a, b, x) keeps its value).  FP32 (the mixed path): one fused
// multiply-add, a single rounding and one instruction instead of two.
__device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// G . r of one face: g0 r0 + g1 r1 (Limiter.cpp:262, Flux.cpp:83-87).
template <typename T>
__device__ __forceinline__ T dot_r(T g0, T g1, T r0, T r1) { return madd(g0, r0, g1 * r1); }

// Flux::EulerFlux, src/Base/Flux.cpp:89-97, given ir = 1 / U[0] (the value
// Rusanov already holds for the same state).
template <typename T>
__device__ __forceinline__ void euler_flux(T gma, const T U[NV], T ir, T n0, T n1, T F[NV]) {
    const T gmo = gma - T(1);
    // P = gmo (U3 - 0.5 ir (U1 U1 + U2 U2)); each madd below is one a*b + c of
    // the reference's expression, grouped as it is written there
    const T P = gmo * madd(-(T(0.5) * ir), madd(U[1], U[1], U[2] * U[2]), U[3]);
    const T E = U[3] + P;
    F[0] = madd(n0, U[1], n1 * U[2]);
    F[1] = madd(n0, madd(U[1] * U[1], ir, P), n1 * U[2] * U[1] * ir);
    F[2] = madd(n1, madd(U[2] * U[2], ir, P), n0 * U[1] * U[2] * ir);
    F[3] = madd(n0 * E * U[1], ir, n1 * E * U[2] * ir);
}

// Flux::Rusanov, src/Base/Flux.cpp:52-81.
template <typename T>
__device__ __forceinline__ void rusanov(T gma, const T UL[NV], const T UR[NV], T n0, T n1, T F[NV]) {
    const T gmo = gma - T(1);
    const T irl = recip(UL[0]);
    const T irr = recip(UR[0]);
    const T nvl = madd(n0, UL[1], n1 * UL[2]) * irl;
    const T nvr = madd(n0, UR[1], n1 * UR[2]) * irr;
    const T qql = T(0.5) * madd(UL[1], UL[1], UL[2] * UL[2]) * irl;
    const T qqr = T(0.5) * madd(UR[1], UR[1], UR[2] * UR[2]) * irr;
    const T pl = gmo * (UL[3] - qql);
    const T pr = gmo * (UR[3] - qqr);
    const T al = vsqrt(gma * pl * irl);
    const T ar = vsqrt(gma * pr * irr);
    // S = max(|nvl - al|, |nvr - ar|, |nvl + al|, |nvr + ar|) as the reference
    // writes it; with a >= 0, max(|q - a|, |q + a|) = |q| + a, and rounding is
    // symmetric and monotone, so the rounded values agree as well (a NaN sound
    // speed drops out of both forms, leaving the other side's bound)
    const T S = dmax(fabs(nvl) + al, fabs(nvr) + ar);
    T FL[NV], FR[NV];
    euler_flux(gma, UL, irl, n0, n1, FL);
    euler_flux(gma, UR, irr, n0, n1, FR);
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = T(0.5) * madd(-S, UR[v] - UL[v], FL[v] + FR[v]);
}

// ---------------------------------------------------------------------------
// FP32-mixed flux in difference form.
//
// The mixed path never rounds a cell's state to FP32 as such: every FP32
// quantity is a difference from the state u of the cell that owns the sum.
// Rounding the FP64 state each step perturbs the solution by about 6e-8 per
// step, which the flow carries and amplifies; differences formed in FP64 and
// then rounded carry a relative error of 6e-8 of the difference instead.
//
// For a face with outward unit normal n, face states u + dO (own side) and
// u + dN (neighbour side), the Rusanov flux (src/Base/Flux.cpp:52-81) less the
// cell's own physical flux E(u).n is
//     F - E(u).n = 1/2 (dE(dO) + dE(dN)) - 1/2 S (dN - dO),
// with dE(d) = E(u + d).n - E(u).n evaluated from d without cancellation and
// S = max |q_side -+ c_side| over both face states (q normal velocity, c sound
// speed).  Summed over a closed cell, sum_f A_f E(u).n_f = E(u).(sum_f A_f n_f)
// vanishes, so the residual is the sum of these differences.
// ---------------------------------------------------------------------------
struct FluxBase {
    float rho, mx, my, e;  // the owning cell's state (FP32)
    float ux, uy, mu, ep;  // velocity, m.u = |m|^2/rho, e + p
    float p;
};

__device__ __forceinline__ FluxBase flux_base(float gma, const double u[NV]) {
    FluxBase b;
    b.rho = (float)u[0];
    b.mx = (float)u[1];
    b.my = (float)u[2];
    b.e = (float)u[3];
    const float ir = recip(b.rho);
    b.ux = b.mx * ir;
    b.uy = b.my * ir;
    b.mu = madd(b.mx, b.ux, b.my * b.uy);
    b.p = (gma - 1.0f) * madd(-0.5f, b.mu, b.e);
    b.ep = b.e + b.p;
    return b;
}

// dE(d) of one face state u + d (see above) and its normal velocity qs and
// sound speed cs.  The normal N = (N0, N1) is the face's A n (length times unit
// normal: the mixed path's geometry planes hold it), so every output is A times
// its unit-normal value; ga2 = gamma |N|^2 = gamma A^2 and q = u.N of the
// owning cell.
__device__ __forceinline__ void euler_flux_delta(float gma, const FluxBase& b, float q, const float d[NV], float n0,
                                                 float n1, float ga2, float dF[NV], float& qs, float& cs) {
    const float irs = recip(b.rho + d[0]);                  // 1 / rho_side
    const float dmn = madd(d[1], n0, d[2] * n1);            // (m_side - m).N
    const float dq = madd(-q, d[0], dmn) * irs;             // q_side - q
    qs = q + dq;
    // |m_side|^2/rho_side - |m|^2/rho = ((2m + dm).dm - (m.u) drho) / rho_side
    const float dmu = madd(madd(2.0f, b.mx, d[1]), d[1], madd(madd(2.0f, b.my, d[2]), d[2], -(b.mu * d[0]))) * irs;
    const float dp = (gma - 1.0f) * madd(-0.5f, dmu, d[3]);  // p_side - p
    cs = vsqrt((b.p + dp) * irs * ga2);                     // A c_side
    dF[0] = dmn;
    dF[1] = madd(d[1], qs, madd(b.mx, dq, dp * n0));
    dF[2] = madd(d[2], qs, madd(b.my, dq, dp * n1));
    dF[3] = madd(d[3] + dp, qs, b.ep * dq);
}

// 2 A (F - E(u).n) of one face (see above) with N = A n_out, the outward frame
// of the owning cell: the factor 1/2 of the Rusanov flux is left to the caller,
// which applies it once to the sum over the cell's faces (exactly: a power of two).
__device__ __forceinline__ void rusanov_delta(float gma, const FluxBase& b, const float dO[NV], const float dN[NV],
                                              float n0, float n1, float F[NV]) {
    const float q = madd(b.ux, n0, b.uy * n1);
    const float ga2 = gma * madd(n0, n0, n1 * n1);
    float FO[NV], FN[NV], qo, co, qn, cn;
    euler_flux_delta(gma, b, q, dO, n0, n1, ga2, FO, qo, co);
    euler_flux_delta(gma, b, q, dN, n0, n1, ga2, FN, qn, cn);
    const float S = dmax(fabsf(qo) + co, fabsf(qn) + cn);  // as in rusanov
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = madd(-S, dN[v] - dO[v], FO[v] + FN[v]);
}

// One face's term of the CFL rate, A (|u.n| + c): FP64 from FP64 geometry (the
// FP64 path), or from the mixed path's FP32 A n with A = |A n|.
__device__ __forceinline__ double cfl_face(double ux, double uy, double c, double A, double n0, double n1) {
    return A * (fabs(ux * n0 + uy * n1) + c);
}
__device__ __forceinline__ double cfl_face(double ux, double uy, double c, float, float N0, float N1) {
    const float A = vsqrt(madd(N0, N0, N1 * N1));
    return fabs(ux * (double)N0 + uy * (double)N1) + c * (double)A;
}

// Limiter face candidate, src/Base/Limiter.cpp:262-266 (eps = 1e-9, :20):
// min(1, (Umax-U)/d) if d > eps, min(1, (Umin-U)/d) if d < -eps, else 1.
// The numerator is selected before the one division, so each face costs one
// divide instead of two; the quotient of the taken branch is the same value.
template <typename T>
__device__ __forceinline__ T limiter_face(T d, T umin_minus_u, T umax_minus_u) {
    const T eps = T(1e-9);
    const bool up = d > eps;
    const bool take = up || (d < -eps);
    const T q = quot(up ? umax_minus_u : umin_minus_u, d);
    return take ? dmin(T(1), q) : T(1);
}

// The limiter of one variable over all faces of a cell, the minimum of
// limiter_face(d_f), from dpos = max(0, max_f d_f) and dneg = min(0, min_f d_f).
// For d > eps the candidate (Umax-U)/d has a numerator >= 0 and falls as d
// grows; for d < -eps, (Umin-U)/d has a numerator <= 0 and falls as |d| grows.
// A correctly rounded quotient keeps that order, so the minimum over the faces
// is the candidate of the largest d and the one of the most negative d: two
// divides per variable instead of one per face, and the same value (faces with
// |d| <= eps, NaN d and NaN quotients drop out as they do face by face).
template <typename T>
__device__ __forceinline__ T limiter_phi(T dpos, T dneg, T umin_minus_u, T umax_minus_u) {
    const T eps = T(1e-9);
    T phi = T(1);
    if (dpos > eps) phi = dmin(phi, quot(umax_minus_u, dpos));
    if (dneg < -eps) phi = dmin(phi, quot(umin_minus_u, dneg));
    return phi;
}

// Element i of a per-cell plane: the plane's base is uniform across the wave
// and i enters as a 32-bit byte offset, so each load is one scalar base plus
// the lane's offset (global_load's saddr form) with no 64-bit vector add per
// plane.  Needs i * sizeof(T) < 2^32: contexts hold at most 2^29 cells.
// The empty asm pins the base in scalar registers and hides how it was formed:
// without it the compiler folds the plane offsets of neighbouring planes into
// 64-bit vector address arithmetic (one v_mad_i64_i32 per plane load).  The
// base must be wave-uniform (the asm reads it from scalar registers).
// Element (plane p, slot k) of cell i in the per-slot geometry arrays (plane
// order GEO_*): blocks of 64 consecutive cells, each holding its GEO_PLANES x
// nslot planes of 64 values back to back.  A wavefront's load of one plane is
// still one contiguous run, and one per-lane address reaches every plane of a
// cell through immediate offsets instead of one scalar base per plane.  The
// arrays hold a whole number of blocks (cell counts padded to 64).
__host__ __device__ __forceinline__ long long geo_index(int nslot, int p, int k, long long i) {
    return ((i >> 6) * (GEO_PLANES * nslot) + (p * nslot + k)) * 64 + (i & 63);
}

template <typename T>
__device__ __forceinline__ T plane_at(const T* __restrict__ base, int i) {
    using GT = __attribute__((address_space(1))) const T;
    using GB = __attribute__((address_space(1))) const char;
    GT* b = (GT*)base;
    asm("" : "+s"(b));
    return *(GT*)((GB*)b + (unsigned)i * (unsigned)sizeof(T));
}

// Cell-state access with an AoSoA tile of width W: component c of cell i of an
// array with NC components lives at (i/W)*NC*W + c*W + i%W.  W == 1 is AoS,
// W >= the cell count behaves as SoA.  The cell's first element is formed in
// 32-bit unsigned arithmetic (NC * cells < 2^32: contexts hold at most 2^29
// cells and NC <= 4), so W a power of two costs a shift and a mask, and c * W
// stays a constant the loads take as an immediate offset.
template <int W, int NC>
__device__ __forceinline__ long long cidx(int i, int c) {
    const unsigned u = (unsigned)i;
    const unsigned x = (u / W) * (NC * W) + u % W;
    return (long long)x + c * W;
}

template <int W, int NC, typename T>
__device__ __forceinline__ void load_cell(const T* __restrict__ a, int i, T out[NC]) {
    if constexpr (W == 1 && sizeof(T) == 8 && NC % 2 == 0) {
        const double2* p = reinterpret_cast<const double2*>(a + (long long)i * NC);
#pragma unroll
        for (int c = 0; c < NC / 2; c++) {
            double2 v = p[c];
            out[2 * c] = v.x;
            out[2 * c + 1] = v.y;
        }
    } else if constexpr (W == 1 && sizeof(T) == 4 && NC % 4 == 0) {
        const float4* p = reinterpret_cast<const float4*>(a + (long long)i * NC);
#pragma unroll
        for (int c = 0; c < NC / 4; c++) {
            float4 v = p[c];
            out[4 * c] = v.x;
            out[4 * c + 1] = v.y;
            out[4 * c + 2] = v.z;
            out[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NC; c++) out[c] = a[cidx<W, NC>(i, c)];
    }
}

template <int W, int NC, typename T>
__device__ __forceinline__ void store_cell(T* __restrict__ a, int i, const T in[NC]) {
    if constexpr (W == 1 && sizeof(T) == 8 && NC % 2 == 0) {
        double2* p = reinterpret_cast<double2*>(a + (long long)i * NC);
#pragma unroll
        for (int c = 0; c < NC / 2; c++) p[c] = make_double2(in[2 * c], in[2 * c + 1]);
    } else if constexpr (W == 1 && sizeof(T) == 4 && NC % 4 == 0) {
        float4* p = reinterpret_cast<float4*>(a + (long long)i * NC);
#pragma unroll
        for (int c = 0; c < NC / 4; c++) p[c] = make_float4(in[4 * c], in[4 * c + 1], in[4 * c + 2], in[4 * c + 3]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; c++) a[cidx<W, NC>(i, c)] = in[c];
    }
}

// CFL rate of a cell, iv * sum_f A_f (|u.n_f| + c): the per-cell bound of the
// step is 1/rate, so dt = cfl * min_i 1/rate_i = cfl * (1 / max_i rate_i)
// exactly (a correctly rounded 1/x is monotone in x) and the kernels reduce the
// maximum rate, leaving one divide per instance.  NaN (a faulted cell) maps to
// 0, which the maximum ignores, as the minimum of bounds ignores NaN.
__device__ __forceinline__ double cfl_rate(double x) { return x >= 0.0 ? x : 0.0; }

// Wave-level maximum over 64 lanes (non-negative values).
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}

// One step of wave_max_dpp: the maximum of x and the x of the lane that the
// DPP control CTRL names (a lane without a valid source reads 0, which leaves
// the maximum of non-negative values unchanged).
template <int CTRL>
__device__ __forceinline__ double dpp_max_step(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return fmax(x, __longlong_as_double(((long long)hi << 32) | (unsigned)lo));
}

// Wave-level maximum over 64 lanes (non-negative values), every lane active:
// quad permutes [1,0,3,2] and [2,3,0,1] and row rotations by 4 and 8 give each
// lane the maximum of its row of 16, the row broadcasts of lanes 15 and 31
// fold the rows into lane 63, whose value is returned to every lane (a
// wave-uniform, scalar result).  Six steps of two DPP moves and one maximum
// instead of the shuffle form's six index computations and LDS permutes.
__device__ __forceinline__ double wave_max_dpp(double x) {
    x = dpp_max_step<0xB1>(x);   // quad_perm [1,0,3,2]
    x = dpp_max_step<0x4E>(x);   // quad_perm [2,3,0,1]
    x = dpp_max_step<0x124>(x);  // row_ror:4
    x = dpp_max_step<0x128>(x);  // row_ror:8
    x = dpp_max_step<0x142>(x);  // row_bcast:15
    x = dpp_max_step<0x143>(x);  // row_bcast:31
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Block-level maximum; every thread receives the result.  `sh` holds one
// double per wave of the block.
__device__ __forceinline__ double block_max(double x, double* sh) {
    x = wave_max(x);
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) sh[w] = x;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < nw; k++) r = fmax(r, sh[k]);
    return r;
}

// Wave-level minimum over 64 lanes.
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off, 64));
    return x;
}

// Block-level minimum; every thread receives the result.  `sh` holds one
// double per wave of the block.
__device__ __forceinline__ double block_min(double x, double* sh) {
    x = wave_min(x);
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) sh[w] = x;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < nw; k++) r = fmin(r, sh[k]);
    return r;
}

// Instance of interior cell i: from the instance of its 64-cell chunk's first
// cell, a forward scan over the (at most few) instance starts inside the chunk.
__device__ __forceinline__ int instance_from(const int* __restrict__ off, const int* __restrict__ chunk_inst, int i) {
    int b = chunk_inst[i >> 6];
    while (off[b + 1] <= i) b++;
    return b;
}

// Interior-cell range of instance b is [off[b], off[b+1]); returns b for cell i.
__device__ __forceinline__ int instance_of(const int* __restrict__ off, int ninst, int i) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

}  // namespace hcfd
