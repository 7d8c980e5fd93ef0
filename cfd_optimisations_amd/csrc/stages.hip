// stages.hip -- one kernel per reference stage, for drop-in parity with each
// Base method on its own, plus state layout conversions and primitives.
//
// Stage arrays are structure-of-arrays in the internal (renumbered) order:
//   sU, sPhi, sR : [4][NcPad]      sG : [4][2][NcPad]   (Base's grad_u[v][d*Nc+i])
// Every cell, ghosts included, is processed as the reference does.
#include "dispatch.hpp"
#include "passes_impl.hpp"
#include "stages.hpp"

namespace hcfd {

// GaussGreen::compute_gradient, src/Base/GaussGreen.cpp:14-40.
__global__ void k_stage_gradient(StageArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nc) return;
    const long long P = a.NcPad;
    double g[NV][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int k = 0; k < a.maxdeg; k++) {
        const Slot s = a.slots[k * P + i];
        if (s.f < 0) break;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
        const int j = slot_nb(s);
        const double If = fr.If, omI = 1 - If;
        for (int v = 0; v < NV; v++) {
            const double ui = a.U[v * P + i], uj = a.U[v * P + j];
            const double uf = isR ? (uj * If + ui * omI) : (ui * If + uj * omI);
            const double au = uf * fr.A;
            const double x0 = au * fr.n0, x1 = au * fr.n1;
            if (isR) {
                g[v][0] = g[v][0] - x0;
                g[v][1] = g[v][1] - x1;
            } else {
                g[v][0] = g[v][0] + x0;
                g[v][1] = g[v][1] + x1;
            }
        }
    }
    const double iv = a.iv[i];
    for (int v = 0; v < NV; v++) {
        a.G[(2 * v) * P + i] = g[v][0] * iv;
        a.G[(2 * v + 1) * P + i] = g[v][1] * iv;
    }
}

// Limiter::limit_gradient, src/Base/Limiter.cpp:19-77.
__global__ void k_stage_limiter(StageArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nc) return;
    const long long P = a.NcPad;
    for (int v = 0; v < NV; v++) {
        const double u = a.U[v * P + i];
        double umin = u, umax = u;
        for (int k = 0; k < a.maxdeg; k++) {
            const Slot s = a.slots[k * P + i];
            if (s.f < 0) break;
            const double uj = a.U[v * P + slot_nb(s)];
            umin = fmin(umin, uj);
            umax = fmax(umax, uj);
        }
        const double g0 = a.G[(2 * v) * P + i], g1 = a.G[(2 * v + 1) * P + i];
        double phi = 1.0;
        for (int k = 0; k < a.maxdeg; k++) {
            const Slot s = a.slots[k * P + i];
            if (s.f < 0) break;
            const bool isR = (s.f & SLOT_SIDE_R) != 0;
            const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
            const double r0 = isR ? fr.rR0 : fr.rL0, r1 = isR ? fr.rR1 : fr.rL1;
            const double d = g0 * r0 + g1 * r1;
            phi = fmin(phi, limiter_face(d, umin - u, umax - u));
        }
        a.PHI[v * P + i] = phi;
    }
}

// BoundaryCondition::update_grads / update_limiter, src/Base/BoundaryConditions/Base.cpp:29-53.
// Ghost g has exactly one face, whose other cell is the interior cell CL.
__global__ void k_stage_bc_copy(StageArgs a, int grads) {
    const int g = a.Nint + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.Nc) return;
    if (a.only_bc >= 0 && a.ghost_bc[g - a.Nint] != a.only_bc) return;
    const long long P = a.NcPad;
    const int l = slot_nb(a.slots[g]);
    if (grads) {
        for (int q = 0; q < 2 * NV; q++) a.G[q * P + g] = -a.G[q * P + l];
    } else {
        for (int v = 0; v < NV; v++) a.PHI[v * P + g] = a.PHI[v * P + l];
    }
}

// Flux::ComputeFlux, src/Base/Flux.cpp:9-50: R accumulates, then R *= iv.
__global__ void k_stage_flux(StageArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nc) return;
    const long long P = a.NcPad;
    double R[NV];
    for (int v = 0; v < NV; v++) R[v] = a.R[v * P + i];
    for (int k = 0; k < a.maxdeg; k++) {
        const Slot s = a.slots[k * P + i];
        if (s.f < 0) break;
        const bool isR = (s.f & SLOT_SIDE_R) != 0;
        const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
        const int L = isR ? slot_nb(s) : i, Rc = isR ? i : slot_nb(s);
        double UfL[NV], UfR[NV];
        for (int v = 0; v < NV; v++) {
            UfL[v] = a.U[v * P + L] + a.PHI[v * P + L] * (a.G[(2 * v) * P + L] * fr.rL0 + a.G[(2 * v + 1) * P + L] * fr.rL1);
            UfR[v] = a.U[v * P + Rc] + a.PHI[v * P + Rc] * (a.G[(2 * v) * P + Rc] * fr.rR0 + a.G[(2 * v + 1) * P + Rc] * fr.rR1);
        }
        double F[NV];
        rusanov(a.gamma, UfL, UfR, fr.n0, fr.n1, F);
        for (int v = 0; v < NV; v++) {
            if (isR) R[v] = R[v] + fr.A * F[v];
            else R[v] = R[v] - fr.A * F[v];
        }
    }
    const double iv = a.iv[i];
    for (int v = 0; v < NV; v++) a.R[v * P + i] = R[v] * iv;
}

// Euler::add_residuals + clear, src/Base/Euler.cpp:107-113 and :93-95.
__global__ void k_stage_update(StageArgs a, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nc) return;
    const long long P = a.NcPad;
    for (int v = 0; v < NV; v++) {
        a.U[v * P + i] += dt * a.R[v * P + i];
        a.R[v * P + i] = 0.0;
    }
}

// BC update_cells on the stage arrays.
__global__ void k_stage_bc_cells(StageArgs a) {
    const int g = a.Nint + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.Nc) return;
    if (a.only_bc >= 0 && a.ghost_bc[g - a.Nint] != a.only_bc) return;
    const long long P = a.NcPad;
    const Slot s = a.slots[g];
    const FaceRec64 fr = a.faces[s.f & (SLOT_SIDE_R - 1)];
    double ul[NV], ug[NV];
    for (int v = 0; v < NV; v++) ul[v] = a.U[v * P + slot_nb(s)];
    ghost_state(a.bctab[a.ghost_bc[g - a.Nint]], a.gamma, ul, fr.n0, fr.n1, ug);
    for (int v = 0; v < NV; v++) a.U[v * P + g] = ug[v];
}

// ---- layout conversions ---------------------------------------------------------

// Caller-order SoA ([4][n] with row pitch `pitch`) -> internal order, layout W (NC comps).
template <int W, int NC>
__global__ void k_import(int n, const int* __restrict__ old_of_new, const double* __restrict__ src, long long pitch,
                         double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = old_of_new[i];
    double x[NC];
    for (int c = 0; c < NC; c++) x[c] = src[c * pitch + o];
    store_cell<W, NC>(dst, i, x);
}

template <int W, int NC>
__global__ void k_export(int n, const int* __restrict__ old_of_new, const double* __restrict__ src,
                         double* __restrict__ dst, long long pitch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[NC];
    load_cell<W, NC>(src, i, x);
    const int o = old_of_new[i];
    for (int c = 0; c < NC; c++) dst[c * pitch + o] = x[c];
}

// Euler::compute_prims, src/Base/Euler.cpp:115-136, into caller-order SoA.
template <int W>
__global__ void k_prims(int n, const int* __restrict__ old_of_new, const double* __restrict__ U, double gma,
                        double* __restrict__ Q, long long pitch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[NV];
    load_cell<W, NV>(U, i, u);
    const double q0 = u[0];
    const double q1 = u[1] / u[0];
    const double q2 = u[2] / u[0];
    const double gmo = gma - 1;
    const double q3 = gmo * (u[3] - 0.5 * q0 * (q1 * q1 + q2 * q2));
    const int o = old_of_new[i];
    Q[o] = q0;
    Q[pitch + o] = q1;
    Q[2 * pitch + o] = q2;
    Q[3 * pitch + o] = q3;
}

// Euler::initialise (src/Base/Euler.cpp:62-77) per cell from primitives in caller order.
template <int W>
__global__ void k_init_prims(int n, const int* __restrict__ old_of_new, const double* __restrict__ Q, long long pitch,
                             double gma, double* __restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = old_of_new[i];
    const double q[NV] = {Q[o], Q[pitch + o], Q[2 * pitch + o], Q[3 * pitch + o]};
    double u[NV];
    u[0] = q[0];
    for (int v = 1; v < NV; v++) u[v] = q[0] * q[v];
    u[3] = q[3] / (gma - 1) + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
    store_cell<W, NV>(U, i, u);
}

template <int W>
__global__ void k_init_uniform(int n, const double* __restrict__ uic, double* __restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[NV] = {uic[0], uic[1], uic[2], uic[3]};
    store_cell<W, NV>(U, i, u);
}

// Per-instance uniform state (batches): cell i of instance inst gets uic[4*inst..].
template <int W>
__global__ void k_init_instances(int Nint, int Nc, const int* __restrict__ inst_off, int ninst,
                                 const int* __restrict__ ghost_inst, const double* __restrict__ uic,
                                 double* __restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    const int b = (i < Nint) ? instance_of(inst_off, ninst, i) : ghost_inst[i - Nint];
    double u[NV] = {uic[4 * b], uic[4 * b + 1], uic[4 * b + 2], uic[4 * b + 3]};
    store_cell<W, NV>(U, i, u);
}

// Ghost states from the interior neighbours on a layout-W state (initialise's BC pass).
template <int W>
__global__ void k_bc_cells_state(int Nint, int Nc, long long NcPad, const Slot* __restrict__ slots,
                                 const FaceRec64* __restrict__ faces, const BCEntry* __restrict__ bctab,
                                 const int* __restrict__ ghost_bc, double gma, double* __restrict__ U) {
    const int g = Nint + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Nc) return;
    const Slot s = slots[g];
    const FaceRec64 fr = faces[s.f & (SLOT_SIDE_R - 1)];
    double ul[NV], ug[NV];
    load_cell<W, NV>(U, slot_nb(s), ul);
    ghost_state(bctab[ghost_bc[g - Nint]], gma, ul, fr.n0, fr.n1, ug);
    store_cell<W, NV>(U, g, ug);
}

// Plain streaming copy, one 16-byte element per lane (one pass over the grid).
__global__ void k_copy16(long long n16, const double2* __restrict__ src, double2* __restrict__ dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

hipError_t launch_copy16(long long n16, const void* src, void* dst, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, n16, (const double2*)src, (double2*)dst);
    return hipGetLastError();
}

// ---- launchers ------------------------------------------------------------------------

static inline int nb(long long n, int b = 256) { return (int)((n + b - 1) / b); }

hipError_t stage_launch(int which, const StageArgs& a, double dt, hipStream_t st) {
    switch (which) {
        case STAGE_GRADIENT: hipLaunchKernelGGL(k_stage_gradient, dim3(nb(a.Nc)), dim3(256), 0, st, a); break;
        case STAGE_LIMITER: hipLaunchKernelGGL(k_stage_limiter, dim3(nb(a.Nc)), dim3(256), 0, st, a); break;
        case STAGE_BC_GRADS:
            if (a.Nc > a.Nint) hipLaunchKernelGGL(k_stage_bc_copy, dim3(nb(a.Nc - a.Nint)), dim3(256), 0, st, a, 1);
            break;
        case STAGE_BC_LIMITER:
            if (a.Nc > a.Nint) hipLaunchKernelGGL(k_stage_bc_copy, dim3(nb(a.Nc - a.Nint)), dim3(256), 0, st, a, 0);
            break;
        case STAGE_FLUX: hipLaunchKernelGGL(k_stage_flux, dim3(nb(a.Nc)), dim3(256), 0, st, a); break;
        case STAGE_UPDATE: hipLaunchKernelGGL(k_stage_update, dim3(nb(a.Nc)), dim3(256), 0, st, a, dt); break;
        case STAGE_BC_CELLS:
            if (a.Nc > a.Nint) hipLaunchKernelGGL(k_stage_bc_cells, dim3(nb(a.Nc - a.Nint)), dim3(256), 0, st, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_import(int width, int ncomp, int n, const int* old_of_new, const double* src, long long pitch,
                         double* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        if (ncomp == NV)
            hipLaunchKernelGGL((k_import<decltype(W)::value, NV>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new, src,
                               pitch, dst);
        else if (ncomp == 2 * NV)
            hipLaunchKernelGGL((k_import<decltype(W)::value, 2 * NV>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new,
                               src, pitch, dst);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    });
}

hipError_t launch_export(int width, int ncomp, int n, const int* old_of_new, const double* src, double* dst,
                         long long pitch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        if (ncomp == NV)
            hipLaunchKernelGGL((k_export<decltype(W)::value, NV>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new, src,
                               dst, pitch);
        else if (ncomp == 2 * NV)
            hipLaunchKernelGGL((k_export<decltype(W)::value, 2 * NV>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new,
                               src, dst, pitch);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    });
}

hipError_t launch_prims(int width, int n, const int* old_of_new, const double* U, double gma, double* Q,
                        long long pitch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        hipLaunchKernelGGL((k_prims<decltype(W)::value>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new, U, gma, Q,
                           pitch);
        return hipGetLastError();
    });
}

hipError_t launch_init_prims(int width, int n, const int* old_of_new, const double* Q, long long pitch, double gma,
                             double* U, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        hipLaunchKernelGGL((k_init_prims<decltype(W)::value>), dim3(nb(n)), dim3(256), 0, st, n, old_of_new, Q, pitch,
                           gma, U);
        return hipGetLastError();
    });
}

hipError_t launch_init_uniform(int width, int n, const double* uic, double* U, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        hipLaunchKernelGGL((k_init_uniform<decltype(W)::value>), dim3(nb(n)), dim3(256), 0, st, n, uic, U);
        return hipGetLastError();
    });
}

hipError_t launch_init_instances(int width, int Nint, int Nc, const int* inst_off, int ninst, const int* ghost_inst,
                                 const double* uic, double* U, hipStream_t st) {
    if (Nc <= 0) return hipSuccess;
    return with_width(width, [&](auto W) {
        hipLaunchKernelGGL((k_init_instances<decltype(W)::value>), dim3(nb(Nc)), dim3(256), 0, st, Nint, Nc, inst_off,
                           ninst, ghost_inst, uic, U);
        return hipGetLastError();
    });
}

hipError_t launch_bc_cells_state(int width, int Nint, int Nc, long long NcPad, const Slot* slots,
                                 const FaceRec64* faces, const BCEntry* bctab, const int* ghost_bc, double gma,
                                 double* U, hipStream_t st) {
    if (Nc <= Nint) return hipSuccess;
    return with_width(width, [&](auto W) {
        hipLaunchKernelGGL((k_bc_cells_state<decltype(W)::value>), dim3(nb(Nc - Nint)), dim3(256), 0, st, Nint, Nc,
                           NcPad, slots, faces, bctab, ghost_bc, gma, U);
        return hipGetLastError();
    });
}

// Caller-order SoA <-> internal-order SoA (stage arrays), ncomp rows.
__global__ void k_soa_gather(int n, int ncomp, const int* __restrict__ old_of_new, const double* __restrict__ src,
                             long long spitch, double* __restrict__ dst, long long dpitch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = old_of_new[i];
    for (int c = 0; c < ncomp; c++) dst[c * dpitch + i] = src[c * spitch + o];
}

__global__ void k_soa_scatter(int n, int ncomp, const int* __restrict__ old_of_new, const double* __restrict__ src,
                              long long spitch, double* __restrict__ dst, long long dpitch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = old_of_new[i];
    for (int c = 0; c < ncomp; c++) dst[c * dpitch + o] = src[c * spitch + i];
}

hipError_t launch_soa_gather(int n, int ncomp, const int* old_of_new, const double* src, long long spitch, double* dst,
                             long long dpitch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soa_gather, dim3(nb(n)), dim3(256), 0, st, n, ncomp, old_of_new, src, spitch, dst, dpitch);
    return hipGetLastError();
}

hipError_t launch_soa_scatter(int n, int ncomp, const int* old_of_new, const double* src, long long spitch, double* dst,
                              long long dpitch, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_soa_scatter, dim3(nb(n)), dim3(256), 0, st, n, ncomp, old_of_new, src, spitch, dst, dpitch);
    return hipGetLastError();
}

}  // namespace hcfd
