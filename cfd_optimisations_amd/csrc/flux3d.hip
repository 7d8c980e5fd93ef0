// flux3d.hip -- the reference's NDIMS == 3 branches on the GPU (FP64, the
// reference's expression order, -ffp-contract=off): Flux::Rusanov with the 3-D
// Flux::EulerFlux (src/Base/Flux.cpp:52-81, 89-92, 98-104), Euler::compute_prims
// (src/Base/Euler.cpp:115-136) and Euler::initialise's state conversion
// (src/Base/Euler.cpp:62-68) for five variables (rho, rho u, rho v, rho w, E).
// The .hmsh format stores two nodes per face and the reference's Domain builds
// two-component normals from them, so no 3-D mesh exists to step; these entry
// points evaluate the 3-D arithmetic on arrays of states (one lane per face or
// state, AoS [n][5] / [n][3] arrays read as contiguous runs).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"
#include "hcfd.h"

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) hcfd::fail(HCFD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace hcfd {
namespace {

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void euler_flux3(double gma, const double U[5], const double n[3], double F[5]) {
    const double gmo = gma - 1;
    const double ir = 1.0 / U[0];
    const double P = gmo * (U[4] - 0.5 * ir * dot3(&U[1], &U[1]));
    F[0] = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
    F[1] = n[0] * (U[1] * U[1] * ir + P) + n[1] * U[2] * U[1] * ir + n[2] * U[3] * U[1] * ir;
    F[2] = n[0] * U[1] * U[2] * ir + n[1] * (U[2] * U[2] * ir + P) + n[2] * U[3] * U[2] * ir;
    F[3] = n[0] * U[1] * U[3] * ir + n[1] * U[2] * U[3] * ir + n[2] * (U[3] * U[3] * ir + P);
    F[4] = n[0] * (U[4] + P) * U[1] * ir + n[1] * (U[4] + P) * U[2] * ir + n[2] * (U[4] + P) * U[3] * ir;
}

__global__ void k_rusanov3d(int n, double gma, const double* __restrict__ UL, const double* __restrict__ UR,
                            const double* __restrict__ nrm, double* __restrict__ F) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    double ul[5], ur[5], nn[3];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        ul[v] = UL[5 * (size_t)f + v];
        ur[v] = UR[5 * (size_t)f + v];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) nn[d] = nrm[3 * (size_t)f + d];
    const double gmo = gma - 1;
    const double irl = 1.0 / ul[0];
    const double irr = 1.0 / ur[0];
    const double nvl = dot3(nn, &ul[1]) * irl;
    const double nvr = dot3(nn, &ur[1]) * irr;
    const double qql = 0.5 * dot3(&ul[1], &ul[1]) * irl;
    const double qqr = 0.5 * dot3(&ur[1], &ur[1]) * irr;
    const double pl = gmo * (ul[4] - qql);
    const double pr = gmo * (ur[4] - qqr);
    const double al = sqrt(gma * pl * irl);
    const double ar = sqrt(gma * pr * irr);
    const double S1 = fmax(fabs(nvl - al), fabs(nvr - ar));
    const double S2 = fmax(fabs(nvl + al), fabs(nvr + ar));
    const double S = fmax(S1, S2);
    double FL[5], FR[5];
    euler_flux3(gma, ul, nn, FL);
    euler_flux3(gma, ur, nn, FR);
#pragma unroll
    for (int v = 0; v < 5; v++) F[5 * (size_t)f + v] = 0.5 * ((FL[v] + FR[v]) - S * (ur[v] - ul[v]));
}

__global__ void k_prims3d(int n, double gma, const double* __restrict__ U, double* __restrict__ Q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[5], q[5];
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = U[5 * (size_t)i + v];
    q[0] = u[0];
#pragma unroll
    for (int v = 1; v < 4; v++) q[v] = u[v] / u[0];
    const double w[3] = {q[1], q[2], q[3]};
    q[4] = (gma - 1) * (u[4] - 0.5 * q[0] * dot3(w, w));
#pragma unroll
    for (int v = 0; v < 5; v++) Q[5 * (size_t)i + v] = q[v];
}

__global__ void k_cons3d(int n, double gma, const double* __restrict__ Q, double* __restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double q[5], u[5];
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = Q[5 * (size_t)i + v];
    u[0] = q[0];
#pragma unroll
    for (int v = 1; v < 5; v++) u[v] = q[0] * q[v];
    u[4] = q[4] / (gma - 1) + 0.5 * q[0] * dot3(&q[1], &q[1]);
#pragma unroll
    for (int v = 0; v < 5; v++) U[5 * (size_t)i + v] = u[v];
}

// Device buffers of one call, freed on every exit path.
struct DevBufs {
    std::vector<double*> p;
    double* get(size_t n) {
        double* x = nullptr;
        const hipError_t e = hipMalloc(&x, sizeof(double) * (n ? n : 1));
        if (e != hipSuccess) fail(HCFD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        p.push_back(x);
        return x;
    }
    ~DevBufs() {
        for (double* x : p) (void)hipFree(x);
    }
};

template <typename F>
int guard3(F&& f) {
    try {
        f();
        return HCFD_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HCFD_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(std::string("unexpected error: ") + e.what());
        return HCFD_ERR_ARG;
    }
}

void check_n(int n, double gma) {
    if (n < 0 || n > (1 << 28)) fail(HCFD_ERR_ARG, "n out of range [0, 2^28]");
    if (!(gma > 1.0)) fail(HCFD_ERR_ARG, "gamma must exceed 1");
}

void h2d(double* d, const double* h, size_t n) {
    if (n) HIPCHK(hipMemcpy(d, h, sizeof(double) * n, hipMemcpyHostToDevice));
}
void d2h(double* h, const double* d, size_t n) {
    if (n) HIPCHK(hipMemcpy(h, d, sizeof(double) * n, hipMemcpyDeviceToHost));
}

}  // namespace
}  // namespace hcfd

using namespace hcfd;

extern "C" {

int hcfd_rusanov3d(int n, double gamma, const double* UL, const double* UR, const double* nrm, double* F) {
    return guard3([&] {
        check_n(n, gamma);
        if (n == 0) return;
        if (!UL || !UR || !nrm || !F) fail(HCFD_ERR_ARG, "null array");
        DevBufs b;
        double *dl = b.get(5 * (size_t)n), *dr = b.get(5 * (size_t)n), *dn = b.get(3 * (size_t)n),
               *df = b.get(5 * (size_t)n);
        h2d(dl, UL, 5 * (size_t)n);
        h2d(dr, UR, 5 * (size_t)n);
        h2d(dn, nrm, 3 * (size_t)n);
        k_rusanov3d<<<(n + 255) / 256, 256>>>(n, gamma, dl, dr, dn, df);
        HIPCHK(hipGetLastError());
        d2h(F, df, 5 * (size_t)n);
    });
}

int hcfd_prims3d(int n, double gamma, const double* U, double* Q) {
    return guard3([&] {
        check_n(n, gamma);
        if (n == 0) return;
        if (!U || !Q) fail(HCFD_ERR_ARG, "null array");
        DevBufs b;
        double *du = b.get(5 * (size_t)n), *dq = b.get(5 * (size_t)n);
        h2d(du, U, 5 * (size_t)n);
        k_prims3d<<<(n + 255) / 256, 256>>>(n, gamma, du, dq);
        HIPCHK(hipGetLastError());
        d2h(Q, dq, 5 * (size_t)n);
    });
}

int hcfd_cons3d(int n, double gamma, const double* Q, double* U) {
    return guard3([&] {
        check_n(n, gamma);
        if (n == 0) return;
        if (!U || !Q) fail(HCFD_ERR_ARG, "null array");
        DevBufs b;
        double *dq = b.get(5 * (size_t)n), *du = b.get(5 * (size_t)n);
        h2d(dq, Q, 5 * (size_t)n);
        k_cons3d<<<(n + 255) / 256, 256>>>(n, gamma, dq, du);
        HIPCHK(hipGetLastError());
        d2h(U, du, 5 * (size_t)n);
    });
}

}  // extern "C"
