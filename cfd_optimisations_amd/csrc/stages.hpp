// stages.hpp -- per-stage kernels (drop-in parity) and state conversions.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace hcfd {

enum StageId {
    STAGE_GRADIENT = 0,
    STAGE_LIMITER,
    STAGE_BC_GRADS,
    STAGE_BC_LIMITER,
    STAGE_FLUX,
    STAGE_UPDATE,
    STAGE_BC_CELLS
};

struct StageArgs {
    int Nc = 0, Nint = 0, NcPad = 0, maxdeg = 0;
    double gamma = 1.4;
    const Slot* slots = nullptr;
    const FaceRec64* faces = nullptr;
    const double* iv = nullptr;
    const BCEntry* bctab = nullptr;
    const int* ghost_bc = nullptr;
    double* U = nullptr;    // [4][NcPad]
    double* G = nullptr;    // [4][2][NcPad]
    double* PHI = nullptr;  // [4][NcPad]
    double* R = nullptr;    // [4][NcPad]
    int only_bc = -1;       // BC stages: restrict to one BC table entry (-1 = all)
};

hipError_t stage_launch(int which, const StageArgs& a, double dt, hipStream_t st);
hipError_t launch_import(int width, int ncomp, int n, const int* old_of_new, const double* src, long long pitch,
                         double* dst, hipStream_t st);
hipError_t launch_export(int width, int ncomp, int n, const int* old_of_new, const double* src, double* dst,
                         long long pitch, hipStream_t st);
hipError_t launch_prims(int width, int n, const int* old_of_new, const double* U, double gma, double* Q,
                        long long pitch, hipStream_t st);
hipError_t launch_init_prims(int width, int n, const int* old_of_new, const double* Q, long long pitch, double gma,
                             double* U, hipStream_t st);
hipError_t launch_init_uniform(int width, int n, const double* uic, double* U, hipStream_t st);
hipError_t launch_init_instances(int width, int Nint, int Nc, const int* inst_off, int ninst, const int* ghost_inst,
                                 const double* uic, double* U, hipStream_t st);
hipError_t launch_bc_cells_state(int width, int Nint, int Nc, long long NcPad, const Slot* slots,
                                 const FaceRec64* faces, const BCEntry* bctab, const int* ghost_bc, double gma,
                                 double* U, hipStream_t st);

hipError_t launch_copy16(long long n16, const void* src, void* dst, int grid, hipStream_t st);
hipError_t launch_soa_gather(int n, int ncomp, const int* old_of_new, const double* src, long long spitch, double* dst,
                             long long dpitch, hipStream_t st);
hipError_t launch_soa_scatter(int n, int ncomp, const int* old_of_new, const double* src, long long spitch, double* dst,
                              long long dpitch, hipStream_t st);

}  // namespace hcfd
