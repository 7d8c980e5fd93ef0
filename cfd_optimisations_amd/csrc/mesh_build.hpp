// mesh_build.hpp -- device mesh construction interface (implemented in mesh_build.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace hcfd {

constexpr int MAX_DEGREE = 8;  // faces per cell supported by the kernels

// Interior cell orders.
enum { ORDER_CALLER = 0, ORDER_RCM = 1, ORDER_SFC = 2 };

// Host view of the (possibly concatenated) input mesh in caller numbering.
struct BuildInput {
    int Nc = 0, Nf = 0, Nn = 0, ninst = 1;
    const int* fc = nullptr;       // [2*Nf]
    const int* fn = nullptr;       // [2*Nf]
    const double* nodes = nullptr; // [2*Nn]
    const int* cn_off = nullptr;   // [Nc+1]
    const int* cn_con = nullptr;
    const int* bcid = nullptr;     // [Nf] BC table entry of a boundary face, -1 for interior faces
    const unsigned char* ghost = nullptr;  // [Nc]
    const int* inst_cell_off = nullptr;    // [ninst+1] caller cell ranges of the instances
    int order = 1;  // ORDER_*: interior cell order
};

// Device-resident products of the build.  Device pointers are owned by the
// context that requested the build and freed with it.
struct DeviceMesh {
    int Nc = 0, Nf = 0, Nint = 0, Nghost = 0, NcPad = 0, maxdeg = 0, ninst = 1;
    int nslot = 3;                  // degree_bucket(maxdeg): slot planes of every per-slot array
    int ncolours = 0;
    int bandwidth_in = 0, bandwidth_out = 0;
    Slot* slots = nullptr;          // [nslot][NcPad], new numbering
    FaceRec64* faces = nullptr;     // [Nf], new numbering
    int* cnb = nullptr;             // [nslot][NcPad] neighbour | twin slot << 29, SLOT_EMPTY
    int* cside = nullptr;           // [NcPad] bit k: the cell is the R side of its slot-k face
    double* cgeo = nullptr;         // per-slot geometry, GEO_PLANES x nslot planes in 64-cell blocks (geo_index)
    double* iv = nullptr;           // [NcPad], new numbering, ghosts 0
    int* ghost_bc = nullptr;        // [max(Nghost,1)] BC table entry per ghost
    int* inst_int_off = nullptr;    // [ninst+1] interior ranges, new numbering
    int* chunk_inst = nullptr;      // [ceil(Nint/64)] instance of cell 64c (batches: a start for inst_of)
    double* xf = nullptr;           // [2*Nf] caller order
    double* xc = nullptr;           // [2*Nc] caller order
    std::vector<int> new_of_old, old_of_new;     // cells
    std::vector<int> face_new_of_old;            // faces
    std::vector<int> colour;                     // caller face order (empty until colour_faces)
    std::vector<int> cf_off, cf_con;             // caller numbering, ascending face id
    std::vector<int> inst_int_off_h;             // [ninst+1]
    std::vector<int> inst_ghost_count;           // [ninst]
};

// Builds everything on the device on `stream` (synchronising as needed).
// Throws hcfd::Error on invalid meshes or HIP failures.
void build_device_mesh(const BuildInput& in, DeviceMesh& out, hipStream_t stream);
void free_device_mesh(DeviceMesh& m);
// Conflict-free face colouring (no two faces of a colour share a cell), caller
// face order, into m.colour / m.ncolours; built once, on the first request.
void colour_faces(DeviceMesh& m, hipStream_t stream);
// dst[i] = (float)src[i], i < n (FP32 geometry and 1/V of the mixed path).
hipError_t launch_d2f(long long n, const double* src, float* dst, hipStream_t st);
// FP32 geometry planes of the mixed path from the FP64 ones (normal planes: A n).
hipError_t launch_geo_mixed(int nslot, long long P, const double* cgeo, float* cgeo32, hipStream_t st);

}  // namespace hcfd
