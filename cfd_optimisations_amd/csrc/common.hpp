// common.hpp -- types and helpers shared by the host code and the gfx950 kernels.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "hcfd.h"

namespace hcfd {

constexpr int NV = 4;  // conserved variables (rho, rho u, rho v, rho E)
constexpr int ND = 2;  // dimensions
constexpr int WAVE = 64;
constexpr int NXCD = 8;          // MI355X: 8 accelerator complex dies, one L2 each
constexpr int PAD = 64;          // cell arrays are padded to a multiple of this
constexpr int SLOT_EMPTY = -1;
constexpr int SLOT_SIDE_R = 1 << 30;  // slot face field: face id | SLOT_SIDE_R when the cell is the face's R side

// Internal error type; converted to a status code at the C boundary.
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void set_last_error(const std::string& msg);

// Face geometry record, FP64 path (64 bytes, one aligned half cache line).
// Values are exactly those of Base::Domain (src/Base/Domain.cpp:73-146):
// A = edge length, n = unit normal oriented L->R, rL/rR = face centre minus
// cell centre (rR = -rL on BC faces), If = interpolation weight of U_L.
struct alignas(64) FaceRec64 {
    double A, n0, n1, rL0, rL1, rR0, rR1, If;
};
static_assert(sizeof(FaceRec64) == 64, "FaceRec64 must be 64 bytes");

// Per-slot geometry planes of the fused passes (one plane per quantity and
// slot, stored in blocks of 64 cells: plane p of slot k of cell i lives at
// ((i / 64) * GEO_PLANES * nslot + p * nslot + k) * 64 + i % 64, geo_index in
// device_math.hpp).  n is the face normal oriented OUT of cell i
// (the L->R normal, negated when i is the face's R cell); r is the face centre
// minus the centre of cell i (rL or rR); If is the face's weight of U_L.
enum { GEO_A = 0, GEO_N0, GEO_N1, GEO_IF, GEO_R0, GEO_R1, GEO_PLANES };

// Boundary-condition table entry (one per BC ordinal per instance).
struct BCEntry {
    int type;      // hcfd_bctype
    int pad;
    double UBC[NV];  // BoundaryCondition::init's UBC (src/Base/BoundaryConditions/Base.cpp:18-22)
    double QBC[NV];
};

// Slot (cell -> face incidence): f = face id | side bit (or SLOT_EMPTY);
// nb = neighbour cell | (slot of this face in the neighbour's list) << 29.
struct Slot {
    int f;
    int nb;
};
constexpr int SLOT_TWIN_SHIFT = 29;
constexpr int SLOT_NB_MASK = (1 << SLOT_TWIN_SHIFT) - 1;

// Slot count the kernels are compiled for, from the largest cell degree:
// 3 (triangles), 4 (quadrilaterals), 8 (polygons up to octagons); -1 above.
inline int degree_bucket(int maxdeg) {
    if (maxdeg <= 3) return 3;
    if (maxdeg <= 4) return 4;
    if (maxdeg <= 8) return 8;
    return -1;
}

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline long long round_up_ll(long long x, long long m) { return (x + m - 1) / m * m; }

// UBC of BoundaryCondition::init, src/Base/BoundaryConditions/Base.cpp:18-22.
inline void cons_of_prims(double gma, const double Q[NV], double U[NV]) {
    U[0] = Q[0];
    for (int v = 1; v < NV; v++) U[v] = Q[0] * Q[v];
    U[3] = Q[3] / (gma - 1) + 0.5 * Q[0] * (Q[1] * Q[1] + Q[2] * Q[2]);
}

}  // namespace hcfd
