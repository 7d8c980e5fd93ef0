// mesh_build.hip -- device-side mesh services run once per context:
//   * cell->face CSR in ascending face id (Mesh::compute_cf_con, src/Mesh.cpp:88-117);
//   * face and cell geometry with Base::Domain's formulas (src/Base/Domain.cpp:37-154);
//   * reverse Cuthill-McKee renumbering of the interior cells by a
//     level-synchronous BFS (one launch per stage per level, all instances of a
//     batch at once), ghosts placed after all interior cells;
//   * faces renumbered by their first visiting cell;
//   * Jones-Plassmann colouring of the face conflict graph (no two faces of a
//     colour share a cell), the working counterpart of the empty
//     Mesh::colour_faces (src/Mesh.cpp:119-121).
// The host orchestration (build_device_mesh) is at the end of this file.
#include <hip/hip_runtime.h>
#include <climits>

#include "device_math.hpp"
#include "mesh_build.hpp"

namespace hcfd {

// ---- connectivity -----------------------------------------------------------

__global__ void k_face_degree(int Nf, const int2* __restrict__ fc, int* __restrict__ deg) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf) return;
    int2 c = fc[f];
    atomicAdd(&deg[c.x], 1);
    atomicAdd(&deg[c.y], 1);
}

__global__ void k_fill_cf(int Nf, const int2* __restrict__ fc, const int* __restrict__ off, int* __restrict__ cursor,
                          int* __restrict__ cf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf) return;
    int2 c = fc[f];
    int pl = atomicAdd(&cursor[c.x], 1);
    cf[off[c.x] + pl] = f;
    int pr = atomicAdd(&cursor[c.y], 1);
    cf[off[c.y] + pr] = f;
}

// Orders each cell's faces by ascending face id (the fill order of
// Mesh::compute_cf_con), which is also the reference's per-cell summation order.
__global__ void k_sort_cf(int Nc, const int* __restrict__ off, int* __restrict__ cf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    int b = off[i], e = off[i + 1];
    for (int a = b + 1; a < e; a++) {
        int key = cf[a];
        int k = a - 1;
        while (k >= b && cf[k] > key) {
            cf[k + 1] = cf[k];
            k--;
        }
        cf[k + 1] = key;
    }
}

// ---- geometry (caller numbering) ---------------------------------------------

// Domain::compute_xf (Domain.cpp:44-59) and compute_nf (:91-116) need xc first;
// xf is independent.
__global__ void k_geom_xf(int Nf, const int2* __restrict__ fn, const double2* __restrict__ nodes,
                          double2* __restrict__ xf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf) return;
    int2 n = fn[f];
    double2 a = nodes[n.x], b = nodes[n.y];
    xf[f] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
}

// Domain::compute_xc (Domain.cpp:61-71): node mean accumulated in cn_con order.
__global__ void k_geom_xc(int Nc, const int* __restrict__ cn_off, const int* __restrict__ cn_con,
                          const double2* __restrict__ nodes, double2* __restrict__ xc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    int b = cn_off[i], e = cn_off[i + 1];
    int nnds = e - b;
    double x = 0.0, y = 0.0;
    for (int k = b; k < e; k++) {
        double2 p = nodes[cn_con[k]];
        x += p.x / nnds;
        y += p.y / nnds;
    }
    xc[i] = make_double2(x, y);
}

// Domain::compute_rf (:73-89), compute_nf (:91-116), compute_If (:118-125).
// bcid[f] >= 0 marks a boundary face (rR = -rL).
__global__ void k_geom_face(int Nf, const int2* __restrict__ fc, const int2* __restrict__ fn,
                            const double2* __restrict__ nodes, const double2* __restrict__ xf,
                            const double2* __restrict__ xc, const int* __restrict__ bcid, FaceRec64* __restrict__ rec) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf) return;
    int2 c = fc[f];
    double2 x = xf[f];
    double2 cl = xc[c.x], cr = xc[c.y];
    double rL0 = x.x - cl.x, rL1 = x.y - cl.y;
    double rR0 = x.x - cr.x, rR1 = x.y - cr.y;
    if (bcid[f] >= 0) {
        rR0 = -rL0;
        rR1 = -rL1;
    }
    int2 n = fn[f];
    double2 a = nodes[n.x], b = nodes[n.y];
    double n0 = (b.y - a.y);
    double n1 = -(b.x - a.x);
    double A = sqrt(n0 * n0 + n1 * n1);
    n0 /= A;
    n1 /= A;
    double ndr = n0 * rL0 + n1 * rL1;
    double sgn = (ndr > 0) ? 1.0 : -1.0;
    n0 = sgn * n0;
    n1 = sgn * n1;
    double nrL = n0 * rL0 + n1 * rL1;
    double nrR = -(n0 * rR0 + n1 * rR1);
    FaceRec64 r;
    r.A = A; r.n0 = n0; r.n1 = n1;
    r.rL0 = rL0; r.rL1 = rL1; r.rR0 = rR0; r.rR1 = rR1;
    r.If = nrL / (nrL + nrR);
    rec[f] = r;
}

// Domain::compute_iv (:127-146) + zero_bc_iv (:148-154).  The reference
// accumulates iv[CL] += A(n.rL)/2 and iv[CR] -= A(n.rR)/2 in face order; each
// cell's faces are walked here in ascending face id, the same order.
__global__ void k_geom_iv(int Nc, const int* __restrict__ off, const int* __restrict__ cf,
                          const int2* __restrict__ fc, const FaceRec64* __restrict__ rec,
                          const unsigned char* __restrict__ ghost, double* __restrict__ iv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    if (ghost[i]) {
        iv[i] = 0.0;
        return;
    }
    double s = 0.0;
    for (int k = off[i]; k < off[i + 1]; k++) {
        int f = cf[k];
        FaceRec64 r = rec[f];
        if (fc[f].x == i) {
            double aL = r.A * (r.n0 * r.rL0 + r.n1 * r.rL1);
            s += aL / 2;
        } else {
            double aR = r.A * (r.n0 * r.rR0 + r.n1 * r.rR1);
            s -= aR / 2;
        }
    }
    iv[i] = 1.0 / s;
}

// ---- reverse Cuthill-McKee ------------------------------------------------------

// Interior-neighbour count of each interior cell (the graph RCM orders).
__global__ void k_interior_degree(int Nc, const int* __restrict__ off, const int* __restrict__ cf,
                                  const int2* __restrict__ fc, const unsigned char* __restrict__ ghost,
                                  int* __restrict__ ideg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    int d = 0;
    if (!ghost[i]) {
        for (int k = off[i]; k < off[i + 1]; k++) {
            int2 c = fc[cf[k]];
            int j = (c.x == i) ? c.y : c.x;
            if (!ghost[j]) d++;
        }
    }
    ideg[i] = d;
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    const unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffu), m, 64);
    const unsigned hi = __shfl_xor((unsigned)(v >> 32), m, 64);
    return ((unsigned long long)hi << 32) | lo;
}

// Per instance: the unplaced interior cell with minimum key (degree, id), or
// (when use_level) minimum (degree, id) among cells on the instance's deepest
// level.  A wave whose cells all belong to one instance reduces first and
// issues one atomic.
__global__ void k_pick_start(int Nc, const int* __restrict__ inst_off, int ninst,
                             const unsigned char* __restrict__ ghost, const int* __restrict__ ideg,
                             const int* __restrict__ cmpos, const int* __restrict__ level,
                             const int* __restrict__ inst_maxlevel, int use_level,
                             unsigned long long* __restrict__ best) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = min(i, Nc - 1);
    const int b = instance_of(inst_off, ninst, ic);
    bool ok = i < Nc && !ghost[i];
    if (ok) ok = use_level ? (level[i] == inst_maxlevel[b]) : (cmpos[i] < 0);
    unsigned long long key = ok ? (((unsigned long long)(unsigned)ideg[i] << 32) | (unsigned)i) : ~0ull;
    int bmin = b, bmax = b;
    for (int m = 32; m > 0; m >>= 1) {
        bmin = min(bmin, __shfl_xor(bmin, m, 64));
        bmax = max(bmax, __shfl_xor(bmax, m, 64));
    }
    if (bmin == bmax) {
        for (int m = 32; m > 0; m >>= 1) {
            const unsigned long long o = shfl_xor_u64(key, m);
            key = o < key ? o : key;
        }
        if ((threadIdx.x & (WAVE - 1)) == 0 && key != ~0ull) atomicMin(&best[b], key);
    } else if (ok) {
        atomicMin(&best[b], key);
    }
}

__global__ void k_max_level(int Nc, const int* __restrict__ inst_off, int ninst, const unsigned char* __restrict__ ghost,
                            const int* __restrict__ level, int* __restrict__ inst_maxlevel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = min(i, Nc - 1);
    const int b = instance_of(inst_off, ninst, ic);
    const bool ok = i < Nc && !ghost[i];
    int lv = ok ? level[i] : -1;
    int bmin = b, bmax = b;
    for (int m = 32; m > 0; m >>= 1) {
        bmin = min(bmin, __shfl_xor(bmin, m, 64));
        bmax = max(bmax, __shfl_xor(bmax, m, 64));
    }
    if (bmin == bmax) {
        for (int m = 32; m > 0; m >>= 1) lv = max(lv, __shfl_xor(lv, m, 64));
        if ((threadIdx.x & (WAVE - 1)) == 0 && lv >= 0) atomicMax(&inst_maxlevel[b], lv);
    } else if (ok) {
        atomicMax(&inst_maxlevel[b], lv);
    }
}

// Seeds the frontier with each instance's chosen start (if any) and places it.
__global__ void k_seed(int ninst, const unsigned long long* __restrict__ best, int* __restrict__ frontier,
                       int* __restrict__ nfront, int* __restrict__ cmpos, int* __restrict__ level, int lvl,
                       int* __restrict__ placed) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int n = 0;
    int base = *placed;
    for (int b = 0; b < ninst; b++) {
        unsigned long long k = best[b];
        if (k == ~0ull) continue;
        int i = (int)(k & 0xffffffffu);
        frontier[n] = i;
        cmpos[i] = base + n;
        level[i] = lvl;
        n++;
    }
    *nfront = n;
    *placed = base + n;
}

// Stage 1: every unplaced interior neighbour is claimed by its first frontier parent.
__global__ void k_bfs_claim(int nfr, const int* __restrict__ frontier, const int* __restrict__ off,
                            const int* __restrict__ cf, const int2* __restrict__ fc,
                            const unsigned char* __restrict__ ghost, const int* __restrict__ cmpos,
                            int* __restrict__ claim) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfr) return;
    int p = frontier[k];
    for (int s = off[p]; s < off[p + 1]; s++) {
        int2 c = fc[cf[s]];
        int j = (c.x == p) ? c.y : c.x;
        if (ghost[j] || cmpos[j] >= 0) continue;
        atomicMin(&claim[j], k);
    }
}

// Gathers the distinct children claimed by frontier entry k, ordered by
// (interior degree, id) as Cuthill-McKee orders siblings.
__device__ __forceinline__ int bfs_children(int k, int p, const int* __restrict__ off, const int* __restrict__ cf,
                                            const int2* __restrict__ fc, const unsigned char* __restrict__ ghost,
                                            const int* __restrict__ cmpos, const int* __restrict__ claim,
                                            const int* __restrict__ ideg, int* ch) {
    int n = 0;
    for (int s = off[p]; s < off[p + 1] && n < MAX_DEGREE; s++) {
        int2 c = fc[cf[s]];
        int j = (c.x == p) ? c.y : c.x;
        if (ghost[j] || cmpos[j] >= 0 || claim[j] != k) continue;
        bool dup = false;
        for (int q = 0; q < n; q++) dup |= (ch[q] == j);
        if (!dup) ch[n++] = j;
    }
    for (int a = 1; a < n; a++) {
        int x = ch[a];
        long long kx = ((long long)ideg[x] << 32) | (unsigned)x;
        int b = a - 1;
        while (b >= 0 && (((long long)ideg[ch[b]] << 32) | (unsigned)ch[b]) > kx) {
            ch[b + 1] = ch[b];
            b--;
        }
        ch[b + 1] = x;
    }
    return n;
}

__global__ void k_bfs_count(int nfr, const int* __restrict__ frontier, const int* __restrict__ off,
                            const int* __restrict__ cf, const int2* __restrict__ fc,
                            const unsigned char* __restrict__ ghost, const int* __restrict__ cmpos,
                            const int* __restrict__ claim, const int* __restrict__ ideg, int* __restrict__ cnt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfr) return;
    int ch[MAX_DEGREE];
    cnt[k] = bfs_children(k, frontier[k], off, cf, fc, ghost, cmpos, claim, ideg, ch);
}

// Stage 3 (after an exclusive scan of cnt): children written to the next
// frontier in CM order and given their global CM position.
__global__ void k_bfs_place(int nfr, const int* __restrict__ frontier, const int* __restrict__ off,
                            const int* __restrict__ cf, const int2* __restrict__ fc,
                            const unsigned char* __restrict__ ghost, int* __restrict__ cmpos,
                            const int* __restrict__ claim, const int* __restrict__ ideg,
                            const int* __restrict__ scan, int base, int lvl, int* __restrict__ level,
                            int* __restrict__ next) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfr) return;
    int ch[MAX_DEGREE];
    int n = bfs_children(k, frontier[k], off, cf, fc, ghost, cmpos, claim, ideg, ch);
    int o = scan[k];
    for (int q = 0; q < n; q++) {
        next[o + q] = ch[q];
        cmpos[ch[q]] = base + o + q;
        level[ch[q]] = lvl;
    }
}

// Sort keys: interior cells by (instance, reversed CM position); ghosts after,
// by the new id of their interior neighbour.
__global__ void k_interior_keys(int Nc, const int* __restrict__ inst_off, int ninst,
                                const unsigned char* __restrict__ ghost, const int* __restrict__ cmpos,
                                unsigned long long* __restrict__ key, int* __restrict__ val, int* __restrict__ n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc || ghost[i]) return;
    int b = instance_of(inst_off, ninst, i);
    int slot = atomicAdd(n, 1);
    key[slot] = ((unsigned long long)(unsigned)b << 32) | (0xffffffffu - (unsigned)cmpos[i]);
    val[slot] = i;
}

__global__ void k_identity_keys(int Nc, const unsigned char* __restrict__ ghost, unsigned long long* __restrict__ key,
                                int* __restrict__ val, int* __restrict__ n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc || ghost[i]) return;
    int slot = atomicAdd(n, 1);
    key[slot] = (unsigned long long)(unsigned)i;
    val[slot] = i;
}

// Index of (x, y) along a Hilbert curve of order 16 (the iterative rotate-and-
// reflect construction of Hilbert's curve; 2^16 x 2^16 cells).
__device__ __forceinline__ unsigned hilbert16(unsigned x, unsigned y) {
    const unsigned n = 1u << 16;
    unsigned d = 0;
    for (unsigned s = n >> 1; s > 0; s >>= 1) {
        const unsigned rx = (x & s) ? 1u : 0u;
        const unsigned ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {  // rotate the quadrant
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const unsigned t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// Sort keys for a space-filling-curve order of the interior cells: (instance,
// Hilbert index of the centroid in the instance's bounding box).
__global__ void k_sfc_keys(int Nc, const int* __restrict__ inst_off, int ninst,
                           const unsigned char* __restrict__ ghost, const double2* __restrict__ xc,
                           const double4* __restrict__ bbox, unsigned long long* __restrict__ key,
                           int* __restrict__ val, int* __restrict__ n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc || ghost[i]) return;
    const int b = instance_of(inst_off, ninst, i);
    const double4 bb = bbox[b];  // xmin, ymin, 1/extent, 1/extent (extent = the larger side)
    const double2 p = xc[i];
    double fx = (p.x - bb.x) * bb.z, fy = (p.y - bb.y) * bb.w;
    fx = fmin(fmax(fx, 0.0), 1.0);
    fy = fmin(fmax(fy, 0.0), 1.0);
    const unsigned qx = min((unsigned)(fx * 65536.0), 65535u);
    const unsigned qy = min((unsigned)(fy * 65536.0), 65535u);
    int slot = atomicAdd(n, 1);
    key[slot] = ((unsigned long long)(unsigned)b << 32) | hilbert16(qx, qy);
    val[slot] = i;
}

__global__ void k_scatter_rank(int n, const int* __restrict__ sorted_val, int base, int* __restrict__ new_of_old) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    new_of_old[sorted_val[r]] = base + r;
}

__global__ void k_ghost_keys(int Nf, const int2* __restrict__ fc, const int* __restrict__ bcid,
                             const int* __restrict__ new_of_old, unsigned long long* __restrict__ key,
                             int* __restrict__ val, int* __restrict__ n) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf || bcid[f] < 0) return;
    int2 c = fc[f];
    int slot = atomicAdd(n, 1);
    key[slot] = ((unsigned long long)(unsigned)new_of_old[c.x] << 32) | (unsigned)f;
    val[slot] = c.y;
}

// Faces ordered by their first visiting cell in the new numbering, then by id.
__global__ void k_face_keys(int Nf, const int2* __restrict__ fc, const int* __restrict__ new_of_old,
                            unsigned long long* __restrict__ key, int* __restrict__ val) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf) return;
    int2 c = fc[f];
    int a = new_of_old[c.x], b = new_of_old[c.y];
    int m = a < b ? a : b;
    key[f] = ((unsigned long long)(unsigned)m << 32) | (unsigned)f;
    val[f] = f;
}

// ---- packing into the internal (renumbered) layout -------------------------

__global__ void k_pack_faces(int Nf, const int* __restrict__ face_old_of_new, const FaceRec64* __restrict__ rec_old,
                             FaceRec64* __restrict__ rec_new) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Nf) return;
    rec_new[g] = rec_old[face_old_of_new[g]];
}

// Slots of every cell (new numbering), kept in ascending ORIGINAL face id so
// each cell sums its faces in the reference's order.
__global__ void k_pack_slots(int Nc, int NcPad, int maxdeg, const int* __restrict__ old_of_new,
                             const int* __restrict__ off, const int* __restrict__ cf, const int2* __restrict__ fc,
                             const int* __restrict__ new_of_old, const int* __restrict__ face_new_of_old,
                             Slot* __restrict__ slots) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NcPad) return;
    if (i >= Nc) {
        for (int k = 0; k < maxdeg; k++) slots[(long long)k * NcPad + i] = Slot{SLOT_EMPTY, 0};
        return;
    }
    int o = old_of_new[i];
    int b = off[o], e = off[o + 1];
    for (int k = 0; k < maxdeg; k++) {
        Slot s{SLOT_EMPTY, 0};
        if (b + k < e) {
            int f = cf[b + k];
            int2 c = fc[f];
            bool isR = (c.x != o);
            int j = isR ? c.x : c.y;
            s.f = face_new_of_old[f] | (isR ? SLOT_SIDE_R : 0);
            // position of this face in the neighbour's own (ascending) face list
            int tw = 0;
            for (int q = off[j]; q < off[j + 1]; q++)
                if (cf[q] == f) {
                    tw = q - off[j];
                    break;
                }
            s.nb = new_of_old[j] | (tw << SLOT_TWIN_SHIFT);
        }
        slots[(long long)k * NcPad + i] = s;
    }
}

// Per-slot cell-major geometry of the fused passes (plane layout GEO_*),
// neighbour|twin indices and side bits, from the slot table and the face
// records.  Slots past a cell's degree hold SLOT_EMPTY and zero geometry.
__global__ void k_pack_cellgeo(int NcPad, int nslot, const Slot* __restrict__ slots,
                               const FaceRec64* __restrict__ faces, int* __restrict__ cnb,
                               int* __restrict__ cside, double* __restrict__ cgeo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NcPad) return;
    const long long P = NcPad;
    int side = 0;
    for (int k = 0; k < nslot; k++) {
        const Slot s = slots[k * P + i];
        double g[GEO_PLANES] = {0, 0, 0, 0, 0, 0};
        int nbx = SLOT_EMPTY;
        if (s.f >= 0) {
            const bool isR = (s.f & SLOT_SIDE_R) != 0;
            const FaceRec64 r = faces[s.f & (SLOT_SIDE_R - 1)];
            g[GEO_A] = r.A;
            g[GEO_N0] = isR ? -r.n0 : r.n0;
            g[GEO_N1] = isR ? -r.n1 : r.n1;
            g[GEO_IF] = r.If;
            g[GEO_R0] = isR ? r.rR0 : r.rL0;
            g[GEO_R1] = isR ? r.rR1 : r.rL1;
            nbx = s.nb;
            if (isR) side |= 1 << k;
        }
        cnb[k * P + i] = nbx;
        for (int p = 0; p < GEO_PLANES; p++) cgeo[geo_index(nslot, p, k, i)] = g[p];
    }
    cside[i] = side;
}

__global__ void k_permute_d(int n, const int* __restrict__ old_of_new, const double* __restrict__ src,
                            double* __restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[old_of_new[i]];
}

// FP32 per-slot geometry of the mixed path: the planes of k_pack_cellgeo
// rounded to FP32, except that the normal planes hold A n (face length times the
// outward unit normal, the product taken in FP64 and rounded once).
__global__ void k_geo_mixed(int nslot, long long P, const double* __restrict__ g, float* __restrict__ g32) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)nslot * P) return;
    const int k = (int)(e / P);
    const long long i = e - k * P;
    const long long a = geo_index(nslot, GEO_A, k, i);
    for (int p = 0; p < GEO_PLANES; p++) {
        const long long q = geo_index(nslot, p, k, i);
        g32[q] = (p == GEO_N0 || p == GEO_N1) ? (float)(g[a] * g[q]) : (float)g[q];
    }
}

__global__ void k_d2f(long long n, const double* __restrict__ src, float* __restrict__ dst) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = (float)src[i];
}

// Ghost g (new id Nint + g) -> BC table entry of its face.
__global__ void k_ghost_bc(int Nf, const int2* __restrict__ fc, const int* __restrict__ bcid,
                           const int* __restrict__ new_of_old, int Nint, int* __restrict__ ghost_bc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf || bcid[f] < 0) return;
    ghost_bc[new_of_old[fc[f].y] - Nint] = bcid[f];
}

// ---- face colouring (Jones-Plassmann) ---------------------------------------

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned long long jp_prio(int f) {
    return (splitmix64((unsigned long long)f) & 0xffffffff00000000ull) | (unsigned)f;
}

// One round: an uncoloured face whose priority beats every uncoloured
// conflicting face takes the smallest colour unused by its coloured conflicts.
__global__ void k_colour_round(int Nf, const int2* __restrict__ fc, const int* __restrict__ off,
                               const int* __restrict__ cf, int* __restrict__ colour, int* __restrict__ remaining) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= Nf || colour[f] >= 0) return;
    unsigned long long pf = jp_prio(f);
    unsigned long long used = 0;
    int2 c = fc[f];
    int cells[2] = {c.x, c.y};
    for (int s = 0; s < 2; s++) {
        int i = cells[s];
        for (int k = off[i]; k < off[i + 1]; k++) {
            int g = cf[k];
            if (g == f) continue;
            int cg = colour[g];
            if (cg < 0) {
                if (jp_prio(g) > pf) {
                    atomicAdd(remaining, 1);
                    return;
                }
            } else if (cg < 64) {
                used |= 1ull << cg;
            }
        }
    }
    int col = 0;
    while (col < 64 && ((used >> col) & 1ull)) col++;
    colour[f] = col;
}

}  // namespace hcfd

// ============================================================================
// Host orchestration
// ============================================================================
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>

namespace hcfd {
namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) fail(HCFD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
T* dalloc(size_t n) {
    T* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) fail(HCFD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return p;
}

struct Scratch {
    std::vector<void*> ptrs;
    template <typename T>
    T* get(size_t n) {
        T* p = dalloc<T>(n);
        ptrs.push_back(p);
        return p;
    }
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

template <typename T>
void h2d(T* d, const T* h, size_t n, hipStream_t st) {
    if (n) HIPCHK(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
}
template <typename T>
void d2h(T* h, const T* d, size_t n, hipStream_t st) {
    if (n) HIPCHK(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
}
int read_int(const int* d, hipStream_t st) {
    int v = 0;
    d2h(&v, d, 1, st);
    return v;
}

// Exclusive scan of n ints (out has n+1 entries; out[n] = total).
void exclusive_scan(const int* d_in, int* d_out, int n, hipStream_t st, Scratch& sc) {
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, d_in, d_out, n + 1, st));
    void* tmp = sc.get<char>(bytes);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, d_in, d_out, n + 1, st));
}

void sort_pairs(unsigned long long* keys, int* vals, int n, hipStream_t st, Scratch& sc) {
    if (n <= 1) return;
    unsigned long long* k2 = sc.get<unsigned long long>(n);
    int* v2 = sc.get<int>(n);
    hipcub::DoubleBuffer<unsigned long long> kb(keys, k2);
    hipcub::DoubleBuffer<int> vb(vals, v2);
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, n, 0, 64, st));
    void* tmp = sc.get<char>(bytes);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kb, vb, n, 0, 64, st));
    if (kb.Current() != keys) HIPCHK(hipMemcpyAsync(keys, kb.Current(), sizeof(unsigned long long) * n,
                                                    hipMemcpyDeviceToDevice, st));
    if (vb.Current() != vals) HIPCHK(hipMemcpyAsync(vals, vb.Current(), sizeof(int) * n, hipMemcpyDeviceToDevice, st));
}

// Counts interior cells whose 1/V is not a positive finite number (an inverted
// or degenerate cell: the method's CFL bound and residual scaling need V > 0).
__global__ void k_check_iv(int Nc, const unsigned char* __restrict__ ghost, const double* __restrict__ iv,
                           int* __restrict__ nbad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nc || ghost[i]) return;
    const double x = iv[i];
    if (!(x > 0.0) || !(x < INFINITY)) atomicAdd(nbad, 1);
}

__global__ void k_validate(int Nf, int Nc, const int2* __restrict__ fc, const int* __restrict__ bcid,
                           const unsigned char* __restrict__ ghost, const int* __restrict__ deg, int* __restrict__ err) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < Nf) {
        int2 c = fc[t];
        if (bcid[t] < 0) {
            if (ghost[c.x] || ghost[c.y]) atomicOr(err, 1);
            if (c.x == c.y) atomicOr(err, 2);
        } else if (ghost[c.x]) {
            atomicOr(err, 4);
        }
    }
    if (t < Nc) {
        int d = deg[t];
        if (ghost[t] && d != 1) atomicOr(err, 8);
        if (!ghost[t] && d == 0) atomicOr(err, 16);
        if (d > MAX_DEGREE) atomicOr(err, 32);
    }
}

__global__ void k_fill_int(int n, int* a, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_u64(int n, unsigned long long* a, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_max_reduce(int n, const int* __restrict__ a, int* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, a[i]);
}

}  // namespace

// Jones-Plassmann rounds over the face conflict graph (faces sharing a cell),
// from the cell->face CSR in caller numbering that the build leaves on the host.
void colour_faces(DeviceMesh& m, hipStream_t st) {
    if (!m.colour.empty() || m.Nf == 0) return;
    const int Nc = m.Nc, Nf = m.Nf;
    // the two cells of every face, from the CSR (order within a face is irrelevant)
    std::vector<int2> fc(Nf, make_int2(-1, -1));
    for (int i = 0; i < Nc; i++)
        for (int k = m.cf_off[i]; k < m.cf_off[i + 1]; k++) {
            int2& c = fc[m.cf_con[k]];
            if (c.x < 0) c.x = i; else c.y = i;
        }
    Scratch sc;
    int2* d_fc = sc.get<int2>(Nf);
    int* d_off = sc.get<int>(Nc + 1);
    int* d_cf = sc.get<int>(2 * (size_t)Nf);
    int* d_colour = sc.get<int>(Nf);
    int* d_n = sc.get<int>(1);
    h2d(d_fc, fc.data(), Nf, st);
    h2d(d_off, m.cf_off.data(), Nc + 1, st);
    h2d(d_cf, m.cf_con.data(), 2 * (size_t)Nf, st);
    k_fill_int<<<nblk(Nf), 256, 0, st>>>(Nf, d_colour, -1);
    for (int round = 0;; round++) {
        if (round > 1000) fail(HCFD_ERR_MESH, "face colouring did not converge");
        HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
        k_colour_round<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_off, d_cf, d_colour, d_n);
        if (read_int(d_n, st) == 0) break;
    }
    std::vector<int> col(Nf);
    d2h(col.data(), d_colour, Nf, st);
    int nc = 0;
    for (int f = 0; f < Nf; f++) nc = std::max(nc, col[f] + 1);
    m.colour.swap(col);
    m.ncolours = nc;
}

void free_device_mesh(DeviceMesh& m) {
    for (void* p : {(void*)m.slots, (void*)m.faces, (void*)m.cnb, (void*)m.cside, (void*)m.cgeo, (void*)m.iv,
                    (void*)m.ghost_bc, (void*)m.inst_int_off, (void*)m.chunk_inst, (void*)m.xf, (void*)m.xc})
        if (p) (void)hipFree(p);
    m.slots = nullptr; m.faces = nullptr; m.cnb = nullptr; m.cside = nullptr; m.cgeo = nullptr; m.iv = nullptr;
    m.ghost_bc = nullptr;
    m.inst_int_off = nullptr; m.chunk_inst = nullptr; m.xf = nullptr; m.xc = nullptr;
}

void build_device_mesh(const BuildInput& in, DeviceMesh& out, hipStream_t st) {
    const int Nc = in.Nc, Nf = in.Nf, Nn = in.Nn, ninst = in.ninst;
    Scratch sc;
    out.Nc = Nc;
    out.Nf = Nf;
    out.ninst = ninst;

    // ---- upload ---------------------------------------------------------------
    int2* d_fc = sc.get<int2>(Nf);
    int2* d_fn = sc.get<int2>(Nf);
    double2* d_nodes = sc.get<double2>(Nn);
    int* d_cn_off = sc.get<int>(Nc + 1);
    int* d_cn_con = sc.get<int>(in.cn_off[Nc]);
    int* d_bcid = sc.get<int>(Nf);
    unsigned char* d_ghost = sc.get<unsigned char>(Nc);
    int* d_inst_off = sc.get<int>(ninst + 1);
    h2d(d_fc, reinterpret_cast<const int2*>(in.fc), Nf, st);
    h2d(d_fn, reinterpret_cast<const int2*>(in.fn), Nf, st);
    h2d(d_nodes, reinterpret_cast<const double2*>(in.nodes), Nn, st);
    h2d(d_cn_off, in.cn_off, Nc + 1, st);
    h2d(d_cn_con, in.cn_con, in.cn_off[Nc], st);
    h2d(d_bcid, in.bcid, Nf, st);
    h2d(d_ghost, in.ghost, Nc, st);
    h2d(d_inst_off, in.inst_cell_off, ninst + 1, st);

    // ---- cell -> face CSR --------------------------------------------------------
    int* d_deg = sc.get<int>(Nc + 1);
    HIPCHK(hipMemsetAsync(d_deg, 0, sizeof(int) * (Nc + 1), st));
    k_face_degree<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_deg);
    int* d_err = sc.get<int>(2);
    HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int) * 2, st));
    k_validate<<<nblk(std::max(Nf, Nc)), 256, 0, st>>>(Nf, Nc, d_fc, d_bcid, d_ghost, d_deg, d_err);
    k_max_reduce<<<nblk(Nc), 256, 0, st>>>(Nc, d_deg, d_err + 1);
    int herr[2];
    d2h(herr, d_err, 2, st);
    if (herr[0]) {
        std::string why;
        if (herr[0] & 1) why += " interior face touches a ghost cell;";
        if (herr[0] & 2) why += " face with CL == CR;";
        if (herr[0] & 4) why += " boundary face has a ghost on the left;";
        if (herr[0] & 8) why += " ghost cell with other than one face;";
        if (herr[0] & 16) why += " interior cell without faces;";
        if (herr[0] & 32) why += " cell with more than 8 faces;";
        fail(HCFD_ERR_MESH, "mesh invariant violated:" + why);
    }
    out.maxdeg = herr[1];
    int* d_off = sc.get<int>(Nc + 1);
    exclusive_scan(d_deg, d_off, Nc, st, sc);
    int* d_cursor = sc.get<int>(Nc);
    HIPCHK(hipMemsetAsync(d_cursor, 0, sizeof(int) * Nc, st));
    int* d_cf = sc.get<int>(2 * (size_t)Nf);
    k_fill_cf<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_off, d_cursor, d_cf);
    k_sort_cf<<<nblk(Nc), 256, 0, st>>>(Nc, d_off, d_cf);

    // ---- geometry, caller numbering ----------------------------------------------
    out.xf = dalloc<double>(2 * (size_t)Nf);
    out.xc = dalloc<double>(2 * (size_t)Nc);
    double2* d_xf = reinterpret_cast<double2*>(out.xf);
    double2* d_xc = reinterpret_cast<double2*>(out.xc);
    k_geom_xf<<<nblk(Nf), 256, 0, st>>>(Nf, d_fn, d_nodes, d_xf);
    k_geom_xc<<<nblk(Nc), 256, 0, st>>>(Nc, d_cn_off, d_cn_con, d_nodes, d_xc);
    FaceRec64* d_rec = sc.get<FaceRec64>(Nf);
    k_geom_face<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_fn, d_nodes, d_xf, d_xc, d_bcid, d_rec);
    double* d_iv_old = sc.get<double>(Nc);
    k_geom_iv<<<nblk(Nc), 256, 0, st>>>(Nc, d_off, d_cf, d_fc, d_rec, d_ghost, d_iv_old);
    {
        int* d_bad = sc.get<int>(1);
        HIPCHK(hipMemsetAsync(d_bad, 0, sizeof(int), st));
        k_check_iv<<<nblk(Nc), 256, 0, st>>>(Nc, d_ghost, d_iv_old, d_bad);
        const int nbad = read_int(d_bad, st);
        if (nbad) fail(HCFD_ERR_MESH, std::to_string(nbad) + " interior cell(s) with non-positive or non-finite volume");
    }

    // ---- interior cell order (RCM) ----------------------------------------------
    std::vector<unsigned char> ghost_h(in.ghost, in.ghost + Nc);
    int Nint = 0;
    for (int i = 0; i < Nc; i++) Nint += ghost_h[i] ? 0 : 1;
    const int Nghost = Nc - Nint;
    out.Nint = Nint;
    out.Nghost = Nghost;
    out.NcPad = round_up(std::max(Nc, 1), PAD);

    int* d_cmpos = sc.get<int>(Nc);
    unsigned long long* d_key = sc.get<unsigned long long>(std::max(Nc, Nf));
    int* d_val = sc.get<int>(std::max(Nc, Nf));
    int* d_n = sc.get<int>(1);
    if (in.order == ORDER_RCM && Nint > 0) {
        int* d_level = sc.get<int>(Nc);
        int* d_claim = sc.get<int>(Nc);
        int* d_ideg = sc.get<int>(Nc);
        int* d_frA = sc.get<int>(Nint);
        int* d_frB = sc.get<int>(Nint);
        int* d_cnt = sc.get<int>(Nint + 1);
        int* d_scan = sc.get<int>(Nint + 1);
        unsigned long long* d_best = sc.get<unsigned long long>(ninst);
        int* d_maxlvl = sc.get<int>(ninst);
        int* d_nfr = sc.get<int>(1);
        k_interior_degree<<<nblk(Nc), 256, 0, st>>>(Nc, d_off, d_cf, d_fc, d_ghost, d_ideg);
        for (int pass = 0; pass < 2; pass++) {
            k_fill_u64<<<nblk(ninst), 256, 0, st>>>(ninst, d_best, ~0ull);
            if (pass == 0) {
                k_fill_int<<<nblk(Nc), 256, 0, st>>>(Nc, d_cmpos, -1);
                k_pick_start<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_ideg, d_cmpos, d_level,
                                                       d_maxlvl, 0, d_best);
            } else {
                // George-Liu refinement: restart from a minimum-degree cell of the deepest level.
                k_fill_int<<<nblk(ninst), 256, 0, st>>>(ninst, d_maxlvl, -1);
                k_max_level<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_level, d_maxlvl);
                k_pick_start<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_ideg, d_cmpos, d_level,
                                                       d_maxlvl, 1, d_best);
                k_fill_int<<<nblk(Nc), 256, 0, st>>>(Nc, d_cmpos, -1);
            }
            k_fill_int<<<nblk(Nc), 256, 0, st>>>(Nc, d_level, -1);
            k_fill_int<<<nblk(Nc), 256, 0, st>>>(Nc, d_claim, INT_MAX);
            // d_n holds the running count of placed cells; k_seed reads and advances it.
            int lvl = 0;
            HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
            k_seed<<<1, 1, 0, st>>>(ninst, d_best, d_frA, d_nfr, d_cmpos, d_level, lvl, d_n);
            int nfr = read_int(d_nfr, st);
            int placed = read_int(d_n, st);
            int guard = 0;
            while (placed < Nint || nfr > 0) {
                if (++guard > 4 * Nint + 16) fail(HCFD_ERR_MESH, "RCM did not terminate");
                if (nfr == 0) {
                    // Disconnected remainder: seed each instance's unplaced minimum-degree cell.
                    k_fill_u64<<<nblk(ninst), 256, 0, st>>>(ninst, d_best, ~0ull);
                    k_pick_start<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_ideg, d_cmpos, d_level,
                                                           d_maxlvl, 0, d_best);
                    lvl++;
                    k_seed<<<1, 1, 0, st>>>(ninst, d_best, d_frA, d_nfr, d_cmpos, d_level, lvl, d_n);
                    nfr = read_int(d_nfr, st);
                    placed = read_int(d_n, st);
                    if (nfr == 0) fail(HCFD_ERR_MESH, "RCM: unplaced cells but no start found");
                    continue;
                }
                k_bfs_claim<<<nblk(nfr), 256, 0, st>>>(nfr, d_frA, d_off, d_cf, d_fc, d_ghost, d_cmpos, d_claim);
                k_bfs_count<<<nblk(nfr), 256, 0, st>>>(nfr, d_frA, d_off, d_cf, d_fc, d_ghost, d_cmpos, d_claim,
                                                       d_ideg, d_cnt);
                exclusive_scan(d_cnt, d_scan, nfr, st, sc);
                int total = read_int(d_scan + nfr, st);
                lvl++;
                k_bfs_place<<<nblk(nfr), 256, 0, st>>>(nfr, d_frA, d_off, d_cf, d_fc, d_ghost, d_cmpos, d_claim,
                                                       d_ideg, d_scan, placed, lvl, d_level, d_frB);
                placed += total;
                HIPCHK(hipMemcpyAsync(d_n, &placed, sizeof(int), hipMemcpyHostToDevice, st));
                std::swap(d_frA, d_frB);
                nfr = total;
            }
            if (placed != Nint) fail(HCFD_ERR_MESH, "RCM placed a wrong number of cells");
        }
        HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
        k_interior_keys<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_cmpos, d_key, d_val, d_n);
    } else if (in.order == ORDER_SFC && Nint > 0) {
        // bounding box of each instance's nodes
        std::vector<double> bb(4 * (size_t)ninst);
        for (int b = 0; b < ninst; b++) {
            double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
            for (int i = in.inst_cell_off[b]; i < in.inst_cell_off[b + 1]; i++)
                for (int q = in.cn_off[i]; q < in.cn_off[i + 1]; q++) {
                    const double x = in.nodes[2 * (size_t)in.cn_con[q]], y = in.nodes[2 * (size_t)in.cn_con[q] + 1];
                    x0 = std::min(x0, x); x1 = std::max(x1, x);
                    y0 = std::min(y0, y); y1 = std::max(y1, y);
                }
            // one scale for both axes: a run of consecutive curve indices then
            // covers a square region (compact tiles) whatever the aspect ratio
            const double ext = std::max(x1 - x0, y1 - y0);
            bb[4 * b] = x0;
            bb[4 * b + 1] = y0;
            bb[4 * b + 2] = ext > 0 ? 1.0 / ext : 1.0;
            bb[4 * b + 3] = ext > 0 ? 1.0 / ext : 1.0;
        }
        double* d_bb = sc.get<double>(bb.size());
        h2d(d_bb, bb.data(), bb.size(), st);
        HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
        k_sfc_keys<<<nblk(Nc), 256, 0, st>>>(Nc, d_inst_off, ninst, d_ghost, d_xc, reinterpret_cast<double4*>(d_bb),
                                              d_key, d_val, d_n);
    } else {
        HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
        k_identity_keys<<<nblk(Nc), 256, 0, st>>>(Nc, d_ghost, d_key, d_val, d_n);
    }
    if (read_int(d_n, st) != Nint) fail(HCFD_ERR_MESH, "interior key count mismatch");
    sort_pairs(d_key, d_val, Nint, st, sc);
    int* d_new_of_old = sc.get<int>(Nc);
    k_scatter_rank<<<nblk(Nint), 256, 0, st>>>(Nint, d_val, 0, d_new_of_old);
    // ghosts after the interior cells, ordered by their interior neighbour
    HIPCHK(hipMemsetAsync(d_n, 0, sizeof(int), st));
    k_ghost_keys<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_bcid, d_new_of_old, d_key, d_val, d_n);
    if (read_int(d_n, st) != Nghost) fail(HCFD_ERR_MESH, "ghost count mismatch (a ghost shared by two faces?)");
    sort_pairs(d_key, d_val, Nghost, st, sc);
    k_scatter_rank<<<nblk(Nghost), 256, 0, st>>>(Nghost, d_val, Nint, d_new_of_old);

    // ---- faces ordered by first visiting cell -------------------------------------------
    k_face_keys<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_new_of_old, d_key, d_val);
    sort_pairs(d_key, d_val, Nf, st, sc);  // d_val = face_old_of_new
    int* d_face_new_of_old = sc.get<int>(Nf);
    k_scatter_rank<<<nblk(Nf), 256, 0, st>>>(Nf, d_val, 0, d_face_new_of_old);

    // ---- pack --------------------------------------------------------------------------
    out.faces = dalloc<FaceRec64>(Nf);
    k_pack_faces<<<nblk(Nf), 256, 0, st>>>(Nf, d_val, d_rec, out.faces);
    int* d_old_of_new = sc.get<int>(Nc);
    // old_of_new from new_of_old
    {
        std::vector<int> no(Nc);
        d2h(no.data(), d_new_of_old, Nc, st);
        out.new_of_old = no;
        out.old_of_new.assign(Nc, -1);
        for (int i = 0; i < Nc; i++) {
            int n = no[i];
            if (n < 0 || n >= Nc || out.old_of_new[n] != -1) fail(HCFD_ERR_MESH, "renumbering is not a permutation");
            out.old_of_new[n] = i;
        }
        h2d(d_old_of_new, out.old_of_new.data(), Nc, st);
    }
    if (Nc > SLOT_NB_MASK) fail(HCFD_ERR_ARG, "more than 2^29 cells");
    out.nslot = degree_bucket(out.maxdeg);
    if (out.nslot < 0) fail(HCFD_ERR_MESH, "cells with more than 8 faces are not supported");
    out.slots = dalloc<Slot>((size_t)out.nslot * out.NcPad);
    k_pack_slots<<<nblk(out.NcPad), 256, 0, st>>>(Nc, out.NcPad, out.nslot, d_old_of_new, d_off, d_cf, d_fc,
                                                  d_new_of_old, d_face_new_of_old, out.slots);
    out.cnb = dalloc<int>((size_t)out.nslot * out.NcPad);
    out.cside = dalloc<int>(out.NcPad);
    out.cgeo = dalloc<double>((size_t)GEO_PLANES * out.nslot * out.NcPad);
    k_pack_cellgeo<<<nblk(out.NcPad), 256, 0, st>>>(out.NcPad, out.nslot, out.slots, out.faces, out.cnb, out.cside,
                                                    out.cgeo);
    out.iv = dalloc<double>(out.NcPad);
    HIPCHK(hipMemsetAsync(out.iv, 0, sizeof(double) * out.NcPad, st));
    k_permute_d<<<nblk(Nc), 256, 0, st>>>(Nc, d_old_of_new, d_iv_old, out.iv);
    out.ghost_bc = dalloc<int>(std::max(Nghost, 1));
    k_ghost_bc<<<nblk(Nf), 256, 0, st>>>(Nf, d_fc, d_bcid, d_new_of_old, Nint, out.ghost_bc);

    // per-instance interior ranges in the new numbering (instances stay contiguous)
    out.inst_int_off_h.assign(ninst + 1, 0);
    out.inst_ghost_count.assign(ninst, 0);
    for (int b = 0; b < ninst; b++) {
        int n = 0, g = 0;
        for (int i = in.inst_cell_off[b]; i < in.inst_cell_off[b + 1]; i++) (ghost_h[i] ? g : n)++;
        out.inst_int_off_h[b + 1] = out.inst_int_off_h[b] + n;
        out.inst_ghost_count[b] = g;
    }
    out.inst_int_off = dalloc<int>(ninst + 1);
    h2d(out.inst_int_off, out.inst_int_off_h.data(), ninst + 1, st);
    {
        // instance of the first cell of every 64-cell chunk: kernels find a cell's
        // instance from there with a short forward scan instead of a binary search
        const int nch = std::max((out.inst_int_off_h[ninst] + 63) / 64, 1);
        std::vector<int> ch(nch, 0);
        int b = 0;
        for (int c = 0; c < nch; c++) {
            while (b + 1 < ninst && out.inst_int_off_h[b + 1] <= 64 * c) b++;
            ch[c] = b;
        }
        out.chunk_inst = dalloc<int>(nch);
        h2d(out.chunk_inst, ch.data(), nch, st);
        HIPCHK(hipStreamSynchronize(st));
    }

    // ---- host copies of the bookkeeping ------------------------------------------------
    // (the face colouring is built on request, colour_faces)
    out.face_new_of_old.resize(Nf);
    d2h(out.face_new_of_old.data(), d_face_new_of_old, Nf, st);
    out.colour.clear();
    out.ncolours = 0;
    out.cf_off.resize(Nc + 1);
    d2h(out.cf_off.data(), d_off, Nc + 1, st);
    out.cf_con.resize(2 * (size_t)Nf);
    d2h(out.cf_con.data(), d_cf, 2 * (size_t)Nf, st);
    out.bandwidth_in = out.bandwidth_out = 0;
    for (int f = 0; f < Nf; f++) {
        int a = in.fc[2 * f], b = in.fc[2 * f + 1];
        if (in.bcid[f] >= 0) continue;
        out.bandwidth_in = std::max(out.bandwidth_in, std::abs(a - b));
        out.bandwidth_out = std::max(out.bandwidth_out, std::abs(out.new_of_old[a] - out.new_of_old[b]));
    }
    HIPCHK(hipStreamSynchronize(st));
}

hipError_t launch_geo_mixed(int nslot, long long P, const double* cgeo, float* cgeo32, hipStream_t st) {
    const long long n = (long long)nslot * P;
    if (n > 0) k_geo_mixed<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(nslot, P, cgeo, cgeo32);
    return hipGetLastError();
}

hipError_t launch_d2f(long long n, const double* src, float* dst, hipStream_t st) {
    if (n > 0) k_d2f<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, src, dst);
    return hipGetLastError();
}

}  // namespace hcfd
