// api.hip -- implementation of the C ABI declared in include/hcfd.h.
//
// A context (hcfd_ctx) is one or more mesh instances concatenated into a
// single renumbered address space on one GPU, plus the state buffers of the
// fused step.  See passes_impl.hpp for the method and mesh_build.hip for the
// device-side mesh construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh_build.hpp"
#include "halo.hpp"
#include "passes.hpp"
#include "stages.hpp"

namespace hcfd {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) fail(HCFD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static T* dmalloc(size_t n) {
    T* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) fail(HCFD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return p;
}

template <typename F>
static int guarded(F&& f) {
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
        set_last_error(e.what());
        return HCFD_ERR_ARG;
    }
}

}  // namespace hcfd

using namespace hcfd;

struct hcfd_ctx {
    int precision = HCFD_FP64;
    int flags = 0;
    hipStream_t st = nullptr;
    int device = 0;
    int ncu = 256;
    DeviceMesh m;
    int Nc = 0, Nf = 0, Nint = 0, Nghost = 0, NcPad = 0, maxdeg = 0, deg_bucket = 3, ninst = 1;
    double gamma = 1.4;
    bool bcs_set = false;
    // per instance, caller numbering
    std::vector<int> inst_cell_off, inst_face_off, inst_bc_base, inst_nbc;
    // device
    BCEntry* bctab = nullptr;
    int nbctab = 0;
    int* d_old_of_new = nullptr;
    int* d_ghost_inst = nullptr;
    double* U[3] = {nullptr, nullptr, nullptr};  // states; the third once a multi-stage integrator is chosen
    int cur = 0;
    int stages = 1;          // time integrator: 1 forward Euler, 2 SSP-RK2, 3 SSP-RK3
    double* FS = nullptr;    // FP64 face states, [deg_bucket][2][NcPad] double2 planes
    float* FS32 = nullptr;   // FP32-mixed face states, [deg_bucket][NcPad] float4
    float* cgeo32 = nullptr; // FP32 per-slot geometry planes (mixed path)
    float* iv32 = nullptr;
    double* pmin_first = nullptr;
    double* pmin_last = nullptr;
    unsigned long long* dtacc = nullptr;
    // during hcfd_step_async_noalloc: the same CFL scratch in the static array
    double* scr_first = nullptr;
    double* scr_last = nullptr;
    unsigned long long* scr_acc = nullptr;
    double* dt = nullptr;
    double* tdev = nullptr;
    double* lastdt = nullptr;
    int* fault = nullptr;
    double* staging = nullptr;   // caller-order SoA scratch, 8 * Nc doubles
    double* sU = nullptr;        // stage arrays (lazy)
    double* sG = nullptr;
    double* sPhi = nullptr;
    double* sR = nullptr;
    double* backup = nullptr;    // autotuner state backup
    hcfd_schedule sched{256, 1, 1, 0, 0, 0};
    // single-kernel step over halo tiles
    HaloTables halo;
    bool fused_ok = false;       // the single-kernel step can run on this mesh
    bool fused_default = false;  // ... and is the default schedule (compact tiles)
    size_t fused_lds = 0;
    bool bound_valid = false;  // dt holds the CFL step of the current state (single-kernel path)
    double bound_cfl = 0.0;    // ... for this CFL number
    double t_host = 0.0;
    double fixed_dt_loaded = NAN;
    // output slots: primitives on the device, their pinned host copy (+ the clock)
    hipStream_t cst = nullptr;   // copy stream (default priority)
    double* out_dev[2] = {nullptr, nullptr};
    double* out_host[2] = {nullptr, nullptr};
    hipEvent_t out_ready[2] = {nullptr, nullptr};
    hipEvent_t out_copied[2] = {nullptr, nullptr};
    bool out_pending[2] = {false, false};
    // captured graph of nbuf steps (U returns to the same buffer)
    hipGraphExec_t gexec = nullptr;
    double graph_dt = NAN, graph_cfl = NAN;
    int graph_cur = -1;
    std::string graph_sched;

    ~hcfd_ctx();
};

// LDS allocation unit assumed when sizing a block's share of the 160 KB per CU.
#ifndef HCFD_LDS_GRANULE
#define HCFD_LDS_GRANULE 2048
#endif
// Blocks per CU the single-kernel step's LDS budget is sized for, the waves per
// SIMD its register budget allows on triangles: FP32-mixed five, FP64 four.
#ifndef HCFD_FUSED_MIXED_BLOCKS
#define HCFD_FUSED_MIXED_BLOCKS 5
#endif
#ifndef HCFD_FUSED_FP64_BLOCKS
#define HCFD_FUSED_FP64_BLOCKS 4
#endif

namespace {

void free_ptr(void* p) {
    if (p) (void)hipFree(p);
}

std::string sched_key(const hcfd_schedule& s) {
    char b[96];
    snprintf(b, sizeof b, "%d/%d/%d/%d/%d/%d", s.block, s.lanes_per_cell, s.aosoa_width, s.persistent, s.xcd_remap,
             s.fused);
    return b;
}

void drop_graph(hcfd_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->graph_cur = -1;
}

// ---- input validation (host) -------------------------------------------------

void validate_view(const hcfd_mesh_view* v, int k) {
    auto bad = [&](const std::string& w) { fail(HCFD_ERR_ARG, "mesh " + std::to_string(k) + ": " + w); };
    if (!v) bad("null mesh view");
    if (v->dim != 2) bad("only 2-D meshes are supported (dim=" + std::to_string(v->dim) + ")");
    if (v->Nn <= 0 || v->Nc <= 0 || v->Nf <= 0 || v->Ncol <= 0 || v->NBCcol < 0 || v->NBCcol > v->Ncol)
        bad("bad sizes");
    if (!v->coloff || !v->nodes || !v->fn_con || !v->fc_con || !v->cn_off || !v->cn_con) bad("null array");
    if (v->coloff[0] != 0 || v->coloff[v->Ncol] != v->Nf) bad("coloff must span [0, Nf]");
    for (int c = 0; c < v->Ncol; c++)
        if (v->coloff[c + 1] < v->coloff[c]) bad("coloff not monotone");
    for (long long q = 0; q < 2LL * v->Nf; q++) {
        if (v->fc_con[q] < 0 || v->fc_con[q] >= v->Nc) bad("fc_con index out of range");
        if (v->fn_con[q] < 0 || v->fn_con[q] >= v->Nn) bad("fn_con index out of range");
    }
    if (v->cn_off[0] != 0) bad("cn_off[0] != 0");
    for (int i = 0; i < v->Nc; i++)
        if (v->cn_off[i + 1] < v->cn_off[i] + 1) bad("cell without nodes");
    for (long long q = 0; q < v->cn_off[v->Nc]; q++)
        if (v->cn_con[q] < 0 || v->cn_con[q] >= v->Nn) bad("cn_con index out of range");
}

// ---- creation ---------------------------------------------------------------------

hcfd_ctx* create_ctx(int n, const hcfd_mesh_view* views, int precision, int flags) {
    if (n <= 0) fail(HCFD_ERR_ARG, "no meshes");
    if (precision != HCFD_FP64 && precision != HCFD_MIXED) fail(HCFD_ERR_ARG, "unknown precision");
    if (precision == HCFD_MIXED && (flags & HCFD_FLAG_EXACT)) fail(HCFD_ERR_ARG, "EXACT applies to the FP64 path");
    for (int k = 0; k < n; k++) validate_view(&views[k], k);

    std::unique_ptr<hcfd_ctx> c(new hcfd_ctx());
    c->precision = precision;
    c->flags = flags;
    c->ninst = n;
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));

    // concatenate instances
    long long Nc = 0, Nf = 0, Nn = 0, ncn = 0;
    for (int k = 0; k < n; k++) {
        Nc += views[k].Nc;
        Nf += views[k].Nf;
        Nn += views[k].Nn;
        ncn += views[k].cn_off[views[k].Nc];
    }
    // neighbour indices are packed in 29 bits, and per-cell planes (padded to
    // PAD cells, at most 8 bytes an entry) are addressed with 32-bit byte offsets
    if (Nc > (1LL << 29) - PAD || Nf >= (1LL << 30) || ncn >= (1LL << 31))
        fail(HCFD_ERR_ARG, "mesh too large: at most 2^29 - 64 cells and 2^30 - 1 faces");
    std::vector<int> fc(2 * Nf), fn(2 * Nf), cn_off(Nc + 1), cn_con(ncn), bcid(Nf, -1);
    std::vector<double> nodes(2 * Nn);
    std::vector<unsigned char> ghost(Nc, 0);
    c->inst_cell_off.assign(n + 1, 0);
    c->inst_face_off.assign(n + 1, 0);
    c->inst_bc_base.assign(n + 1, 0);
    c->inst_nbc.assign(n, 0);
    long long co = 0, fo = 0, no = 0, cno = 0;
    int bcbase = 0;
    for (int k = 0; k < n; k++) {
        const hcfd_mesh_view& v = views[k];
        const int F_BC = v.coloff[v.Ncol - v.NBCcol];
        for (long long f = 0; f < v.Nf; f++) {
            fc[2 * (fo + f)] = v.fc_con[2 * f] + (int)co;
            fc[2 * (fo + f) + 1] = v.fc_con[2 * f + 1] + (int)co;
            fn[2 * (fo + f)] = v.fn_con[2 * f] + (int)no;
            fn[2 * (fo + f) + 1] = v.fn_con[2 * f + 1] + (int)no;
        }
        for (int col = v.Ncol - v.NBCcol; col < v.Ncol; col++)
            for (int f = v.coloff[col]; f < v.coloff[col + 1]; f++) {
                bcid[fo + f] = bcbase + (col - (v.Ncol - v.NBCcol));
                ghost[co + v.fc_con[2 * f + 1]] = 1;
            }
        (void)F_BC;
        for (long long q = 0; q < 2LL * v.Nn; q++) nodes[2 * no + q] = v.nodes[q];
        for (int i = 0; i < v.Nc; i++) cn_off[co + i] = v.cn_off[i] + (int)cno;
        for (long long q = 0; q < v.cn_off[v.Nc]; q++) cn_con[cno + q] = v.cn_con[q] + (int)no;
        co += v.Nc;
        fo += v.Nf;
        no += v.Nn;
        cno += v.cn_off[v.Nc];
        c->inst_cell_off[k + 1] = (int)co;
        c->inst_face_off[k + 1] = (int)fo;
        c->inst_nbc[k] = v.NBCcol;
        bcbase += v.NBCcol;
        c->inst_bc_base[k + 1] = bcbase;
    }
    cn_off[Nc] = (int)cno;

    BuildInput in;
    in.Nc = (int)Nc;
    in.Nf = (int)Nf;
    in.Nn = (int)Nn;
    in.ninst = n;
    in.fc = fc.data();
    in.fn = fn.data();
    in.nodes = nodes.data();
    in.cn_off = cn_off.data();
    in.cn_con = cn_con.data();
    in.bcid = bcid.data();
    in.ghost = ghost.data();
    in.inst_cell_off = c->inst_cell_off.data();
    in.order = (flags & HCFD_FLAG_NO_RENUMBER) ? ORDER_CALLER : (flags & HCFD_FLAG_ORDER_SFC) ? ORDER_SFC : ORDER_RCM;
    build_device_mesh(in, c->m, c->st);

    c->Nc = c->m.Nc;
    c->Nf = c->m.Nf;
    c->Nint = c->m.Nint;
    c->Nghost = c->m.Nghost;
    c->NcPad = c->m.NcPad;
    c->maxdeg = c->m.maxdeg;
    c->deg_bucket = c->m.nslot;
    for (int k = 0; k < n; k++)
        if (c->m.inst_int_off_h[k + 1] == c->m.inst_int_off_h[k]) fail(HCFD_ERR_MESH, "instance without interior cells");

    // permutation + ghost instance on the device
    c->d_old_of_new = dmalloc<int>(c->Nc);
    HIPCHK(hipMemcpyAsync(c->d_old_of_new, c->m.old_of_new.data(), sizeof(int) * c->Nc, hipMemcpyHostToDevice, c->st));
    {
        std::vector<int> gi(std::max(c->Nghost, 1), 0);
        for (int k = 0; k < n; k++)
            for (int i = c->inst_cell_off[k]; i < c->inst_cell_off[k + 1]; i++)
                if (ghost[i]) gi[c->m.new_of_old[i] - c->Nint] = k;
        c->d_ghost_inst = dmalloc<int>(gi.size());
        HIPCHK(hipMemcpyAsync(c->d_ghost_inst, gi.data(), sizeof(int) * gi.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }

    // state buffers (sized for the widest tile layout: NcPad is a multiple of 64)
    const size_t P = c->NcPad;
    c->U[0] = dmalloc<double>(4 * P);
    c->U[1] = dmalloc<double>(4 * P);
    HIPCHK(hipMemsetAsync(c->U[0], 0, sizeof(double) * 4 * P, c->st));
    HIPCHK(hipMemsetAsync(c->U[1], 0, sizeof(double) * 4 * P, c->st));
    // face states: deg_bucket slots x 4 values per cell, slot-major planes
    const size_t nfs = P * (size_t)c->deg_bucket * NV;
    if (precision == HCFD_FP64) {
        c->FS = dmalloc<double>(nfs);
    } else {
        c->FS32 = dmalloc<float>(nfs);
        const size_t ngeo = (size_t)GEO_PLANES * c->deg_bucket * P;
        c->cgeo32 = dmalloc<float>(ngeo);
        c->iv32 = dmalloc<float>(P);
        // FP32 geometry from the FP64 planes (same values rounded; normal planes A n)
        HIPCHK(launch_geo_mixed(c->deg_bucket, P, c->m.cgeo, c->cgeo32, c->st));
        HIPCHK(launch_d2f((long long)P, c->m.iv, c->iv32, c->st));
    }
    const int maxtiles = (c->Nint + 15) / 16 + 1;  // smallest pass-A tile: 64 threads / 4 lanes per cell
    c->pmin_first = dmalloc<double>(maxtiles);
    c->pmin_last = dmalloc<double>(maxtiles);
    c->dtacc = dmalloc<unsigned long long>(n + 1);  // + the one-instance reduction's block ticket
    c->dt = dmalloc<double>(n);
    c->tdev = dmalloc<double>(n);
    c->lastdt = dmalloc<double>(n);
    c->fault = dmalloc<int>(n);
    {
        HIPCHK(hipMemsetAsync(c->dtacc, 0, sizeof(unsigned long long) * (n + 1), c->st));  // maximum rate: 0
        HIPCHK(hipMemsetAsync(c->tdev, 0, sizeof(double) * n, c->st));
        HIPCHK(hipMemsetAsync(c->lastdt, 0, sizeof(double) * n, c->st));
        HIPCHK(hipMemsetAsync(c->fault, 0, sizeof(int) * n, c->st));
    }
    c->staging = dmalloc<double>(8 * (size_t)c->Nc);
    c->nbctab = std::max(bcbase, 1);
    c->bctab = dmalloc<BCEntry>(c->nbctab);
    HIPCHK(hipMemsetAsync(c->bctab, 0, sizeof(BCEntry) * c->nbctab, c->st));
    c->sched.xcd_remap = 1;
    // halo tiles of the single-kernel step; used when their LDS footprint fits
    // LDS share of one block at the occupancy the registers allow: FP64 four
    // blocks of four wavefronts per CU on triangles (three on quadrilaterals, two
    // on octagon meshes), FP32-mixed five on triangles (four otherwise)
    {
        const int blocks = precision == HCFD_MIXED ? (c->deg_bucket <= 3 ? HCFD_FUSED_MIXED_BLOCKS : 4)
                                                   : (c->deg_bucket <= 3 ? HCFD_FUSED_FP64_BLOCKS
                                                                         : c->deg_bucket <= 4 ? 3 : 2);
        // the share is rounded down to whole HCFD_LDS_GRANULE units: measured, blocks
        // of 53.9 KB (under a third of 160 KB) run two per CU, blocks of 53.5 KB three
        const size_t budget = (size_t)160 * 1024 / blocks / HCFD_LDS_GRANULE * HCFD_LDS_GRANULE;
        build_halo_tables(c->Nint, c->NcPad, c->deg_bucket, c->m.cnb, c->m.cside,
                          fused_cells(precision == HCFD_MIXED), precision == HCFD_MIXED, budget, c->halo, c->st);
    }
    c->fused_lds = fused_lds_bytes(precision == HCFD_MIXED, c->deg_bucket, c->halo.smax, c->halo.fsmax);
    {
        int lim = 0;
        HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
        // runnable when the largest tile's halo fits in LDS with two blocks per CU
        c->fused_ok = c->fused_lds <= (size_t)std::min(lim, 80 * 1024) && c->halo.lt && c->halo.unresolved == 0;
        // default only for compact tiles (at most 5% of the tiles had to be split):
        // strip-shaped tiles, as a reverse Cuthill-McKee order makes, recompute more
        // halo than they update; Hilbert tiles keep the halo near a quarter
        c->fused_default = c->fused_ok && c->halo.split_tiles * 20 <= c->halo.ntiles;
    }
    if (c->fused_ok) {
        build_ring_inputs(c->deg_bucket, (long long)P, precision == HCFD_MIXED ? nullptr : c->m.cgeo, c->m.iv,
                          c->cgeo32, c->iv32, c->halo, c->st);
        if (n > 1) build_instance_tiles(c->m.inst_int_off, n, c->halo, c->st);
    }
    c->sched.fused = c->fused_default ? 1 : 0;
    HIPCHK(hipStreamSynchronize(c->st));
    return c.release();
}

void fill_bctab(hcfd_ctx* c, int inst, int nbc, const int* types, const double* qbc, double gamma,
                std::vector<BCEntry>& tab) {
    if (nbc != c->inst_nbc[inst])
        fail(HCFD_ERR_ARG, "instance " + std::to_string(inst) + " has " + std::to_string(c->inst_nbc[inst]) +
                               " BC colours, got " + std::to_string(nbc));
    for (int b = 0; b < nbc; b++) {
        if (types[b] < HCFD_SYMMETRY || types[b] > HCFD_SUBOUTLET) fail(HCFD_ERR_ARG, "unknown BC type");
        BCEntry e{};
        e.type = types[b];
        for (int v = 0; v < NV; v++) e.QBC[v] = qbc[NV * b + v];
        cons_of_prims(gamma, e.QBC, e.UBC);
        tab[c->inst_bc_base[inst] + b] = e;
    }
}

void upload_bctab(hcfd_ctx* c, const std::vector<BCEntry>& tab) {
    HIPCHK(hipMemcpyAsync(c->bctab, tab.data(), sizeof(BCEntry) * tab.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->bcs_set = true;
    drop_graph(c);
}

void require_bcs(hcfd_ctx* c) {
    if (!c->bcs_set) fail(HCFD_ERR_STATE, "boundary conditions not set (hcfd_set_bcs)");
}

void apply_bc_cells(hcfd_ctx* c, double* U) {
    HIPCHK(launch_bc_cells_state(c->sched.aosoa_width, c->Nint, c->Nc, c->NcPad, c->m.slots, c->m.faces, c->bctab,
                                 c->m.ghost_bc, c->gamma, U, c->st));
}

void reset_time(hcfd_ctx* c) {
    c->t_host = 0.0;
    HIPCHK(hipMemsetAsync(c->tdev, 0, sizeof(double) * c->ninst, c->st));
    HIPCHK(hipMemsetAsync(c->fault, 0, sizeof(int) * c->ninst, c->st));
}

// Caller-order SoA host arrays -> staging (pitch Nc).
void upload_soa(hcfd_ctx* c, const double* const X[4], int ncomp_per_var) {
    for (int v = 0; v < NV; v++) {
        if (!X[v]) fail(HCFD_ERR_ARG, "null array");
        HIPCHK(hipMemcpyAsync(c->staging + (size_t)v * ncomp_per_var * c->Nc, X[v],
                              sizeof(double) * ncomp_per_var * c->Nc, hipMemcpyHostToDevice, c->st));
    }
}

void download_soa(hcfd_ctx* c, double* const X[4], int ncomp_per_var) {
    for (int v = 0; v < NV; v++) {
        if (!X[v]) fail(HCFD_ERR_ARG, "null array");
        HIPCHK(hipMemcpyAsync(X[v], c->staging + (size_t)v * ncomp_per_var * c->Nc,
                              sizeof(double) * ncomp_per_var * c->Nc, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
}

// ---- fused stepping -------------------------------------------------------------------

struct Launch {
    PassVariant v;
    PassArgs a;
    int grid = 0, block = 0;
};

int nbuf(const hcfd_ctx* c) { return c->U[2] ? 3 : 2; }

// pass: 0 = pass A (tiles of block / lanes_per_cell cells), 1 = pass B (one cell per lane).
Launch make_launch(hcfd_ctx* c, bool cfl, int cur, int pass) {
    Launch L;
    const hcfd_schedule& s = c->sched;
    L.v.precision = c->precision;
    L.v.cfl = cfl;
    L.v.width = s.aosoa_width;
    L.v.maxdeg = c->deg_bucket;
    L.v.lpc = pass == 0 ? s.lanes_per_cell : 1;
    const bool fused = pass == 2;
    PassArgs& a = L.a;
    a.Nint = c->Nint;
    a.Nc = c->Nc;
    a.NcPad = c->NcPad;
    a.ninst = c->ninst;
    a.tile_cells = fused ? fused_cells(c->precision == HCFD_MIXED) : s.block / L.v.lpc;
    a.ntiles = fused ? c->halo.ntiles : (c->Nint + a.tile_cells - 1) / a.tile_cells;
    a.tile_start = fused ? c->halo.tile_start : nullptr;
    a.inst_tiles = fused ? c->halo.inst_tiles : nullptr;
    a.xcd_remap = s.xcd_remap;
    a.tiles_per_xcd = (a.ntiles + NXCD - 1) / NXCD;
    a.gamma = c->gamma;
    a.cnb = c->m.cnb;
    a.cside = c->m.cside;
    a.cgeo = c->m.cgeo;
    a.cgeo32 = c->cgeo32;
    a.iv = c->m.iv;
    a.iv32 = c->iv32;
    a.U = c->U[cur];
    a.Unext = c->U[(cur + 1) % nbuf(c)];
    a.FS = c->FS;
    a.FS32 = c->FS32;
    a.inst_off = c->m.inst_int_off;
    a.chunk_inst = c->m.chunk_inst;
    a.pmin_first = c->scr_first ? c->scr_first : c->pmin_first;
    a.pmin_last = c->scr_first ? c->scr_last : c->pmin_last;
    a.dtacc = c->scr_first ? c->scr_acc : c->dtacc;
    a.dt = c->dt;
    a.bctab = c->bctab;
    a.ghost_bc = c->m.ghost_bc;
    a.fault = c->fault;
    a.h1_off = c->halo.h1_off;
    a.h1 = c->halo.h1;
    a.s_off = c->halo.s_off;
    a.s_ngt = c->halo.s_ngt;
    a.s2 = c->halo.s2;
    a.lt = c->halo.lt;
    a.lh = c->halo.lh;
    a.nh1tot = std::max(c->halo.nh1, 1LL);
    a.hgeo = c->halo.hgeo;
    a.hiv = c->halo.hiv;
    a.hgeo32 = c->halo.hgeo32;
    a.hiv32 = c->halo.hiv32;
    a.smax = c->halo.smax;
    a.fsmax = c->halo.fsmax;
    a.tacc = c->tdev;
    a.lastdt = c->lastdt;
    L.block = fused ? 256 : s.block;
    if (s.persistent > 0) {
        int g = std::min(a.ntiles, s.persistent * c->ncu);
        if (s.xcd_remap) g = std::max(NXCD, g / NXCD * NXCD);
        L.grid = std::max(g, 1);
    } else {
        L.grid = s.xcd_remap ? NXCD * a.tiles_per_xcd : a.ntiles;
    }
    return L;
}

void load_fixed_dt(hcfd_ctx* c, double dt) {
    if (c->fixed_dt_loaded == dt) return;
    std::vector<double> h(c->ninst, dt);
    HIPCHK(hipMemcpyAsync(c->dt, h.data(), sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->fixed_dt_loaded = dt;
}

// Single-kernel path, CFL mode: dt must hold the CFL step of the current state
// before the first step (afterwards every step reduces the bound of its result).
// A call with another CFL number than the one the stored dt was reduced with
// takes the bound again, so its first step uses its own CFL number.
void ensure_bound(hcfd_ctx* c, double cfl) {
    if (!c->sched.fused || !(cfl > 0) || (c->bound_valid && c->bound_cfl == cfl)) return;
    Launch L = make_launch(c, true, c->cur, 2);
    HIPCHK(launch_cfl_bound(L.v, L.a, L.grid, c->st));
    HIPCHK(launch_dt_reduce(L.a, cfl, c->dt, c->st));
    c->bound_valid = true;
    c->bound_cfl = cfl;
}

// Buffers of one step from U[cur]: stage k reads U[in[k]] and writes U[out[k]];
// stages after the first combine with U[cur] (U^n) and never overwrite it.
// Forward Euler alternates between the buffers; SSP-RK uses all three.
struct StagePlan {
    int in[3] = {0, 0, 0}, out[3] = {0, 0, 0};
    int result = 0;
};
StagePlan plan_step(const hcfd_ctx* c, int cur) {
    StagePlan p;
    const int nb = nbuf(c), x = (cur + 1) % nb, y = (cur + 2) % nb;
    p.in[0] = cur;
    p.out[0] = x;
    p.result = x;
    if (c->stages >= 2) {
        p.in[1] = x;
        p.out[1] = y;
        p.result = y;
    }
    if (c->stages == 3) {
        p.in[2] = y;
        p.out[2] = x;
        p.result = x;
    }
    return p;
}

// Shu-Osher coefficients of stage k: U_k = RK_A U^n + RK_B (U_{k-1} + dt R(U_{k-1}))
// (Gottlieb, Shu and Tadmor, SIAM Rev. 43 (2001) 89-112, eqs. 3.1-3.2).
const double RK_A[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
const double RK_B[4][3] = {{1, 1, 1}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};

void set_stage(PassArgs& a, const hcfd_ctx* c, const StagePlan& p, int k, int cur) {
    a.U = c->U[p.in[k]];
    a.Unext = c->U[p.out[k]];
    a.U0 = k > 0 ? c->U[cur] : nullptr;
    a.rk_a = RK_A[c->stages][k];
    a.rk_b = RK_B[c->stages][k];
}

// Enqueues one step reading U[cur]; returns the buffer holding U^{n+1}.  Every
// stage uses the dt of U^n.  Events (when given): ev[0] step start, ev[1] after
// the first stage's pass A (or single kernel), ev[2] after its CFL reduction,
// ev[3] step end.
int enqueue_step(hcfd_ctx* c, int cur, double cfl, hipEvent_t* ev = nullptr) {
    const bool use_cfl = cfl > 0;
    const StagePlan p = plan_step(c, cur);
    const int S = c->stages;
    if (ev) HIPCHK(hipEventRecord(ev[0], c->st));
    for (int k = 0; k < S; k++) {
        const bool last = k == S - 1;
        if (c->sched.fused) {
            // the first stage advances the clocks by dt (fixed, or reduced from
            // the CFL bound of U^n); in CFL mode the last stage leaves the bound
            // of U^{n+1} per tile and the reduction gives dt of the next step
            Launch F = make_launch(c, use_cfl && last, cur, 2);
            set_stage(F.a, c, p, k, cur);
            F.a.advance = k == 0;
            HIPCHK(launch_fused_step(F.v, F.a, F.grid, c->fused_lds, c->st));
            if (ev && k == 0) HIPCHK(hipEventRecord(ev[1], c->st));
            if (use_cfl && last) HIPCHK(launch_dt_reduce(F.a, cfl, c->dt, c->st));
            if (ev && k == 0) HIPCHK(hipEventRecord(ev[2], c->st));
        } else {
            // CFL: pass A of the first stage takes the bound of U^n; pass B of
            // the first stage advances the clocks by dt
            Launch L = make_launch(c, use_cfl && k == 0, cur, 0);
            Launch LB = make_launch(c, use_cfl && k == 0, cur, 1);
            set_stage(L.a, c, p, k, cur);
            set_stage(LB.a, c, p, k, cur);
            LB.a.advance = k == 0;
            HIPCHK(launch_pass_a(L.v, L.a, L.grid, L.block, c->st));
            if (ev && k == 0) HIPCHK(hipEventRecord(ev[1], c->st));
            if (use_cfl && k == 0) HIPCHK(launch_dt_reduce(L.a, cfl, c->dt, c->st));
            if (ev && k == 0) HIPCHK(hipEventRecord(ev[2], c->st));
            HIPCHK(launch_pass_b(LB.v, LB.a, LB.grid, LB.block, c->st));
        }
    }
    if (ev) HIPCHK(hipEventRecord(ev[3], c->st));
    c->bound_valid = c->sched.fused && use_cfl;
    c->bound_cfl = cfl;
    return p.result;
}

void ensure_graph(hcfd_ctx* c, double dt, double cfl) {
    const std::string key = sched_key(c->sched);
    if (c->gexec && c->graph_cur == c->cur && c->graph_dt == dt && c->graph_cfl == cfl && c->graph_sched == key)
        return;
    drop_graph(c);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    try {
        int k = c->cur;
        for (int q = 0; q < nbuf(c); q++) k = enqueue_step(c, k, cfl);
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(c->st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    HIPCHK(hipStreamEndCapture(c->st, &g));
    hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    c->graph_cur = c->cur;
    c->graph_dt = dt;
    c->graph_cfl = cfl;
    c->graph_sched = key;
}

void run_steps(hcfd_ctx* c, int nsteps, double dt, double cfl) {
    require_bcs(c);
    if (nsteps < 0) fail(HCFD_ERR_ARG, "nsteps < 0");
    if (!(cfl > 0) && !(dt > 0)) fail(HCFD_ERR_ARG, "need dt > 0 or cfl > 0");
    if (!(cfl > 0)) load_fixed_dt(c, dt);
    else c->fixed_dt_loaded = NAN;  // the CFL reduction overwrites dt
    ensure_bound(c, cfl);
    int done = 0;
    if ((c->flags & HCFD_FLAG_GRAPH) && nsteps >= nbuf(c)) {
        ensure_graph(c, dt, cfl);
        // every step moves cur by the same amount: nbuf steps bring it back
        for (; done + nbuf(c) <= nsteps; done += nbuf(c)) HIPCHK(hipGraphLaunch(c->gexec, c->st));
    }
    for (; done < nsteps; done++) c->cur = enqueue_step(c, c->cur, cfl);
}

// ---- allocation-free stepping ---------------------------------------------------

}  // namespace (reopened below)

// CFL scratch of hcfd_step_async_noalloc: the per-tile partial maxima of the
// CFL rate (tile's first and last instance) and the per-instance atomic maxima
// that hcfd_step_async keeps in buffers the context allocates.  One copy per
// GPU, held until the process ends.  Host code never uses this name for the
// memory: the device address comes from hipGetSymbolAddress (scratch_base).
constexpr size_t HCFD_NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned long long scratch[HCFD_NOALLOC_SCRATCH_BYTES / sizeof(unsigned long long)];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_base[MAX_DEVICES] = {};

// The scratch array's device address on `device` (the current device), looked
// up once per device.
void* scratch_base(int device) {
    if (device < 0 || device >= MAX_DEVICES) fail(HCFD_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_base[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(HCFD_ERR_HIP, "hipGetSymbolAddress returned no address for the CFL scratch");
        g_scratch_base[device] = p;
    }
    return g_scratch_base[device];
}

__global__ void k_zero_u64(int n, unsigned long long* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_fill_f64(int n, double* __restrict__ p, double x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = x;
}

// run_steps with the CFL scratch in the static array, eager launches only (no
// graph), the fixed dt stored by a kernel: nothing is allocated, on the device
// or the host, and nothing waits for the device.
void run_steps_noalloc(hcfd_ctx* c, int nsteps, double dt, double cfl) {
    require_bcs(c);
    if (nsteps < 0) fail(HCFD_ERR_ARG, "nsteps < 0");
    if (!(cfl > 0) && !(dt > 0)) fail(HCFD_ERR_ARG, "need dt > 0 or cfl > 0");
    // tiles of the schedule's CFL partials: the halo tiles of the single-kernel
    // step, or pass A's tiles of block / lanes_per_cell cells
    const long long tiles = c->sched.fused ? (long long)c->halo.ntiles
                                           : (c->Nint + c->sched.block / c->sched.lanes_per_cell - 1) /
                                                 (c->sched.block / c->sched.lanes_per_cell);
    const size_t need = sizeof(double) * (size_t)(2 * tiles + c->ninst + 1);
    if (need > HCFD_NOALLOC_SCRATCH_BYTES)
        fail(HCFD_ERR_ARG, "hcfd_step_async_noalloc: this context's schedule needs " + std::to_string(need) +
                               " bytes of CFL scratch (" + std::to_string(tiles) + " tiles, " +
                               std::to_string(c->ninst) + " instances); the static array holds " +
                               std::to_string(HCFD_NOALLOC_SCRATCH_BYTES) + " (use hcfd_step_async)");
    int dev0 = 0;
    HIPCHK(hipGetDevice(&dev0));
    if (dev0 != c->device) HIPCHK(hipSetDevice(c->device));
    struct Restore {
        hcfd_ctx* c;
        int dev0;
        ~Restore() {
            c->scr_first = c->scr_last = nullptr;
            c->scr_acc = nullptr;
            if (dev0 != c->device) (void)hipSetDevice(dev0);
        }
    } restore{c, dev0};
    double* base = static_cast<double*>(scratch_base(c->device));
    c->scr_first = base;
    c->scr_last = base + tiles;
    c->scr_acc = reinterpret_cast<unsigned long long*>(base + 2 * tiles);
    // the per-instance maxima and the reduction's ticket start at zero (the array
    // keeps the last call's contents)
    const int nb = (c->ninst + 256) / 256;
    k_zero_u64<<<nb, 256, 0, c->st>>>(c->ninst + 1, c->scr_acc);
    HIPCHK(hipGetLastError());
    if (!(cfl > 0)) {
        k_fill_f64<<<nb, 256, 0, c->st>>>(c->ninst, c->dt, dt);
        HIPCHK(hipGetLastError());
        c->fixed_dt_loaded = dt;
    } else {
        c->fixed_dt_loaded = NAN;  // the CFL reduction overwrites dt
    }
    ensure_bound(c, cfl);
    for (int done = 0; done < nsteps; done++) c->cur = enqueue_step(c, c->cur, cfl);
}

double device_time(hcfd_ctx* c) {
    double t = 0;
    HIPCHK(hipMemcpyAsync(&t, c->tdev, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return t;
}

void ensure_stage_arrays(hcfd_ctx* c) {
    if (c->precision != HCFD_FP64) fail(HCFD_ERR_STATE, "stage entry points need an FP64 context");
    if (c->sU) return;
    const size_t P = c->NcPad;
    c->sU = dmalloc<double>(4 * P);
    c->sG = dmalloc<double>(8 * P);
    c->sPhi = dmalloc<double>(4 * P);
    c->sR = dmalloc<double>(4 * P);
    HIPCHK(hipMemsetAsync(c->sU, 0, sizeof(double) * 4 * P, c->st));
    HIPCHK(hipMemsetAsync(c->sG, 0, sizeof(double) * 8 * P, c->st));
    HIPCHK(hipMemsetAsync(c->sPhi, 0, sizeof(double) * 4 * P, c->st));
    HIPCHK(hipMemsetAsync(c->sR, 0, sizeof(double) * 4 * P, c->st));
}

StageArgs stage_args(hcfd_ctx* c) {
    StageArgs a;
    a.Nc = c->Nc;
    a.Nint = c->Nint;
    a.NcPad = c->NcPad;
    a.maxdeg = c->maxdeg;
    a.gamma = c->gamma;
    a.slots = c->m.slots;
    a.faces = c->m.faces;
    a.iv = c->m.iv;
    a.bctab = c->bctab;
    a.ghost_bc = c->m.ghost_bc;
    a.U = c->sU;
    a.G = c->sG;
    a.PHI = c->sPhi;
    a.R = c->sR;
    return a;
}

void run_stage(hcfd_ctx* c, int which, double dt) {
    ensure_stage_arrays(c);
    if (which == STAGE_BC_CELLS) require_bcs(c);
    HIPCHK(stage_launch(which, stage_args(c), dt, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
}

void change_width(hcfd_ctx* c, int neww) {
    const int oldw = c->sched.aosoa_width;
    if (oldw == neww) return;
    // U -> caller-order staging -> new layout (identity permutation keeps internal order)
    std::vector<int> ident(c->Nc);
    for (int i = 0; i < c->Nc; i++) ident[i] = i;
    int* d_ident = dmalloc<int>(c->Nc);
    HIPCHK(hipMemcpyAsync(d_ident, ident.data(), sizeof(int) * c->Nc, hipMemcpyHostToDevice, c->st));
    HIPCHK(launch_export(oldw, NV, c->Nc, d_ident, c->U[c->cur], c->staging, c->Nc, c->st));
    HIPCHK(launch_import(neww, NV, c->Nc, d_ident, c->staging, c->Nc, c->U[c->cur], c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    (void)hipFree(d_ident);
    c->sched.aosoa_width = neww;
    c->bound_valid = false;
    drop_graph(c);
}

void check_schedule(const hcfd_ctx* c, const hcfd_schedule& s) {
    if (s.block != 64 && s.block != 128 && s.block != 256) fail(HCFD_ERR_ARG, "block must be 64, 128 or 256");
    if (s.lanes_per_cell != 1 && s.lanes_per_cell != 4) fail(HCFD_ERR_ARG, "lanes_per_cell must be 1 or 4");
    if (s.aosoa_width != 1 && s.aosoa_width != 8 && s.aosoa_width != 64)
        fail(HCFD_ERR_ARG, "aosoa_width must be 1, 8 or 64");
    if (s.persistent < 0 || s.persistent > 64) fail(HCFD_ERR_ARG, "persistent must be in [0, 64]");
    if (s.fused != 0 && s.fused != 1) fail(HCFD_ERR_ARG, "fused must be 0 or 1");
    if (s.fused && (s.block != 256 || s.lanes_per_cell != 1))
        fail(HCFD_ERR_ARG, "the single-kernel step needs block 256 and one lane per cell");
    if (s.fused && !c->fused_ok) fail(HCFD_ERR_ARG, "this mesh's halo tiles do not fit in LDS (fused unavailable)");
}

}  // namespace

hcfd_ctx::~hcfd_ctx() {
    if (st) (void)hipStreamSynchronize(st);
    if (cst) (void)hipStreamSynchronize(cst);
    for (int k = 0; k < 2; k++) {
        if (out_host[k]) (void)hipHostFree(out_host[k]);
        if (out_ready[k]) (void)hipEventDestroy(out_ready[k]);
        if (out_copied[k]) (void)hipEventDestroy(out_copied[k]);
        free_ptr(out_dev[k]);
    }
    if (cst) (void)hipStreamDestroy(cst);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    free_device_mesh(m);
    free_halo_tables(halo);
    for (void* p : {(void*)bctab, (void*)d_old_of_new, (void*)d_ghost_inst, (void*)U[0], (void*)U[1], (void*)U[2], (void*)FS,
                    (void*)FS32, (void*)cgeo32, (void*)iv32, (void*)pmin_first, (void*)pmin_last,
                    (void*)dtacc, (void*)dt, (void*)tdev, (void*)lastdt, (void*)fault, (void*)staging, (void*)sU,
                    (void*)sG, (void*)sPhi, (void*)sR, (void*)backup})
        free_ptr(p);
    if (st) (void)hipStreamDestroy(st);
}

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int hcfd_version(void) { return 10000; }

int hcfd_device_count(int* n) {
    return guarded([&] {
        if (!n) fail(HCFD_ERR_ARG, "null");
        HIPCHK(hipGetDeviceCount(n));
    });
}

int hcfd_create(const hcfd_mesh_view* mesh, int precision, int flags, hcfd_ctx** out) {
    return guarded([&] {
        if (!out) fail(HCFD_ERR_ARG, "null out");
        *out = create_ctx(1, mesh, precision, flags);
    });
}

int hcfd_batch_create(int n, const hcfd_mesh_view* meshes, const hcfd_case* cases, double gamma, int precision,
                      int flags, hcfd_ctx** out) {
    return guarded([&] {
        if (!out || !meshes || !cases) fail(HCFD_ERR_ARG, "null argument");
        if (!(gamma > 1.0)) fail(HCFD_ERR_ARG, "gamma must exceed 1");
        std::unique_ptr<hcfd_ctx> c(create_ctx(n, meshes, precision, flags));
        c->gamma = gamma;
        c->bound_valid = false;
        std::vector<BCEntry> tab(c->nbctab);
        std::vector<double> uic(4 * (size_t)n);
        for (int k = 0; k < n; k++) {
            fill_bctab(c.get(), k, meshes[k].NBCcol, cases[k].bctype, cases[k].qbc, gamma, tab);
            const double* q = cases[k].qic;
            double u[NV];
            u[0] = q[0];
            for (int v = 1; v < NV; v++) u[v] = q[0] * q[v];
            u[3] = q[3] / (gamma - 1) + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
            for (int v = 0; v < NV; v++) uic[4 * k + v] = u[v];
        }
        upload_bctab(c.get(), tab);
        double* d_uic = dmalloc<double>(uic.size());
        HIPCHK(hipMemcpyAsync(d_uic, uic.data(), sizeof(double) * uic.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(launch_init_instances(c->sched.aosoa_width, c->Nint, c->Nc, c->m.inst_int_off, c->ninst,
                                     c->d_ghost_inst, d_uic, c->U[c->cur], c->st));
        apply_bc_cells(c.get(), c->U[c->cur]);
        reset_time(c.get());
        HIPCHK(hipStreamSynchronize(c->st));
        (void)hipFree(d_uic);
        *out = c.release();
    });
}

int hcfd_destroy(hcfd_ctx* ctx) {
    return guarded([&] { delete ctx; });
}

int hcfd_get_info(hcfd_ctx* c, hcfd_info* o) {
    return guarded([&] {
        if (!c || !o) fail(HCFD_ERR_ARG, "null");
        o->Nc = c->Nc;
        o->Nf = c->Nf;
        o->Nint = c->Nint;
        o->Nghost = c->Nghost;
        o->ninst = c->ninst;
        o->maxdeg = c->maxdeg;
        o->precision = c->precision;
        o->flags = c->flags;
        o->bandwidth_in = c->m.bandwidth_in;
        o->bandwidth_out = c->m.bandwidth_out;
        o->ncolours = c->m.ncolours;
        size_t P = c->NcPad;
        long long b = 0;
        b += (long long)c->deg_bucket * P * (sizeof(Slot) + sizeof(int) + 8 * GEO_PLANES) + P * (8 + 4);
        b += (long long)c->Nf * sizeof(FaceRec64);
        b += 8LL * P * 8;                         // U x2
        b += (c->FS ? 32LL : c->FS32 ? 16LL : 0) * (long long)P * c->deg_bucket;
        b += (c->cgeo32 ? 4LL * GEO_PLANES * c->deg_bucket * P + P * 4 : 0);
        b += 8LL * c->Nc * 8 + 16LL * c->Nf + 16LL * c->Nc;  // staging, xf, xc
        const long long nh1 = std::max(c->halo.nh1, 1LL);       // ring input copy
        if (c->halo.hgeo) b += 8LL * (GEO_PLANES * c->deg_bucket + 1) * nh1;
        if (c->halo.hgeo32) b += 4LL * (GEO_PLANES * c->deg_bucket + 1) * nh1;
        o->device_bytes = b;
        o->sched = c->sched;
        o->fused_available = c->fused_ok ? 1 : 0;
        o->fused_lds_bytes = (int)c->fused_lds;
        o->halo_max_cells = c->halo.smax;
    });
}

int hcfd_get_geometry(hcfd_ctx* c, double* xf, double* xc, double* rf, double* nf, double* Af, double* If,
                      double* iv) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        if (xf) HIPCHK(hipMemcpy(xf, c->m.xf, sizeof(double) * 2 * c->Nf, hipMemcpyDeviceToHost));
        if (xc) HIPCHK(hipMemcpy(xc, c->m.xc, sizeof(double) * 2 * c->Nc, hipMemcpyDeviceToHost));
        if (rf || nf || Af || If) {
            std::vector<FaceRec64> rec(c->Nf);
            HIPCHK(hipMemcpy(rec.data(), c->m.faces, sizeof(FaceRec64) * c->Nf, hipMemcpyDeviceToHost));
            for (int f = 0; f < c->Nf; f++) {
                const FaceRec64& r = rec[c->m.face_new_of_old[f]];
                if (rf) {
                    rf[4 * f] = r.rL0; rf[4 * f + 1] = r.rL1; rf[4 * f + 2] = r.rR0; rf[4 * f + 3] = r.rR1;
                }
                if (nf) { nf[2 * f] = r.n0; nf[2 * f + 1] = r.n1; }
                if (Af) Af[f] = r.A;
                if (If) If[f] = r.If;
            }
        }
        if (iv) {
            std::vector<double> h(c->NcPad);
            HIPCHK(hipMemcpy(h.data(), c->m.iv, sizeof(double) * c->NcPad, hipMemcpyDeviceToHost));
            for (int i = 0; i < c->Nc; i++) iv[i] = h[c->m.new_of_old[i]];
        }
    });
}

int hcfd_get_cf(hcfd_ctx* c, int* cf_off, int* cf_con) {
    return guarded([&] {
        if (!c || !cf_off || !cf_con) fail(HCFD_ERR_ARG, "null");
        std::memcpy(cf_off, c->m.cf_off.data(), sizeof(int) * c->m.cf_off.size());
        std::memcpy(cf_con, c->m.cf_con.data(), sizeof(int) * c->m.cf_con.size());
    });
}

int hcfd_get_permutation(hcfd_ctx* c, int* new_of_old) {
    return guarded([&] {
        if (!c || !new_of_old) fail(HCFD_ERR_ARG, "null");
        std::memcpy(new_of_old, c->m.new_of_old.data(), sizeof(int) * c->Nc);
    });
}

int hcfd_get_face_colours(hcfd_ctx* c, int* colour_of_face, int* ncolours) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        colour_faces(c->m, c->st);  // built on the first request
        if (colour_of_face) std::memcpy(colour_of_face, c->m.colour.data(), sizeof(int) * c->Nf);
        if (ncolours) *ncolours = c->m.ncolours;
    });
}

int hcfd_set_bcs(hcfd_ctx* c, int nbc, const int* bctype, const double* qbc, double gamma) {
    return guarded([&] {
        if (!c || (nbc > 0 && (!bctype || !qbc))) fail(HCFD_ERR_ARG, "null argument");
        if (c->ninst != 1) fail(HCFD_ERR_STATE, "batches take their BCs at creation");
        if (!(gamma > 1.0)) fail(HCFD_ERR_ARG, "gamma must exceed 1");
        c->gamma = gamma;
        c->bound_valid = false;  // the CFL bound depends on gamma
        std::vector<BCEntry> tab(c->nbctab);
        fill_bctab(c, 0, nbc, bctype, qbc, gamma, tab);
        upload_bctab(c, tab);
    });
}

int hcfd_initialise(hcfd_ctx* c, const double qic[4]) {
    return guarded([&] {
        if (!c || !qic) fail(HCFD_ERR_ARG, "null");
        require_bcs(c);
        // Euler::initialise, src/Base/Euler.cpp:62-77
        double u[NV];
        u[0] = qic[0];
        for (int v = 1; v < NV; v++) u[v] = qic[0] * qic[v];
        u[3] = qic[3] / (c->gamma - 1) + 0.5 * qic[0] * (qic[1] * qic[1] + qic[2] * qic[2]);
        HIPCHK(hipMemcpyAsync(c->staging, u, sizeof(u), hipMemcpyHostToDevice, c->st));
        HIPCHK(launch_init_uniform(c->sched.aosoa_width, c->Nc, c->staging, c->U[c->cur], c->st));
        apply_bc_cells(c, c->U[c->cur]);
        reset_time(c);
        c->bound_valid = false;
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_initialise_prims(hcfd_ctx* c, const double* const Q[4]) {
    return guarded([&] {
        if (!c || !Q) fail(HCFD_ERR_ARG, "null");
        require_bcs(c);
        upload_soa(c, Q, 1);
        HIPCHK(launch_init_prims(c->sched.aosoa_width, c->Nc, c->d_old_of_new, c->staging, c->Nc, c->gamma,
                                 c->U[c->cur], c->st));
        apply_bc_cells(c, c->U[c->cur]);
        reset_time(c);
        c->bound_valid = false;
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_set_state(hcfd_ctx* c, const double* const U[4]) {
    return guarded([&] {
        if (!c || !U) fail(HCFD_ERR_ARG, "null");
        upload_soa(c, U, 1);
        HIPCHK(launch_import(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->staging, c->Nc, c->U[c->cur], c->st));
        c->bound_valid = false;
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_get_state(hcfd_ctx* c, double* const U[4]) {
    return guarded([&] {
        if (!c || !U) fail(HCFD_ERR_ARG, "null");
        HIPCHK(launch_export(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->U[c->cur], c->staging, c->Nc, c->st));
        download_soa(c, U, 1);
    });
}

int hcfd_get_prims(hcfd_ctx* c, double* const Q[4]) {
    return guarded([&] {
        if (!c || !Q) fail(HCFD_ERR_ARG, "null");
        HIPCHK(launch_prims(c->sched.aosoa_width, c->Nc, c->d_old_of_new, c->U[c->cur], c->gamma, c->staging, c->Nc,
                            c->st));
        download_soa(c, Q, 1);
    });
}

int hcfd_stage_set(hcfd_ctx* c, int which, const double* const X[4]) {
    return guarded([&] {
        if (!c || !X) fail(HCFD_ERR_ARG, "null");
        ensure_stage_arrays(c);
        const int per = (which == HCFD_ARR_G) ? 2 : 1;
        double* dst = which == HCFD_ARR_U ? c->sU : which == HCFD_ARR_G ? c->sG : which == HCFD_ARR_PHI ? c->sPhi
                    : which == HCFD_ARR_R ? c->sR : nullptr;
        if (!dst) fail(HCFD_ERR_ARG, "unknown array");
        upload_soa(c, X, per);
        HIPCHK(launch_soa_gather(c->Nc, NV * per, c->d_old_of_new, c->staging, c->Nc, dst, c->NcPad, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_stage_get(hcfd_ctx* c, int which, double* const X[4]) {
    return guarded([&] {
        if (!c || !X) fail(HCFD_ERR_ARG, "null");
        ensure_stage_arrays(c);
        const int per = (which == HCFD_ARR_G) ? 2 : 1;
        const double* src = which == HCFD_ARR_U ? c->sU : which == HCFD_ARR_G ? c->sG : which == HCFD_ARR_PHI ? c->sPhi
                          : which == HCFD_ARR_R ? c->sR : nullptr;
        if (!src) fail(HCFD_ERR_ARG, "unknown array");
        HIPCHK(launch_soa_scatter(c->Nc, NV * per, c->d_old_of_new, src, c->NcPad, c->staging, c->Nc, c->st));
        download_soa(c, X, per);
    });
}

int hcfd_stage_gradient(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_GRADIENT, 0); }); }
int hcfd_stage_limiter(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_LIMITER, 0); }); }
int hcfd_stage_bc_grads(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_BC_GRADS, 0); }); }
int hcfd_stage_bc_limiter(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_BC_LIMITER, 0); }); }
int hcfd_stage_flux(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_FLUX, 0); }); }
int hcfd_stage_update(hcfd_ctx* c, double dt) { return guarded([&] { run_stage(c, STAGE_UPDATE, dt); }); }
int hcfd_stage_bc_cells(hcfd_ctx* c) { return guarded([&] { run_stage(c, STAGE_BC_CELLS, 0); }); }

int hcfd_stage_bc(hcfd_ctx* c, int what, int bc) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        if (c->ninst != 1) fail(HCFD_ERR_STATE, "per-BC stages apply to single-mesh contexts");
        if (bc < -1 || bc >= c->inst_nbc[0]) fail(HCFD_ERR_ARG, "BC ordinal out of range");
        const int which = what == 0 ? STAGE_BC_CELLS : what == 1 ? STAGE_BC_GRADS : what == 2 ? STAGE_BC_LIMITER : -1;
        if (which < 0) fail(HCFD_ERR_ARG, "what must be 0 (cells), 1 (gradients) or 2 (limiter)");
        ensure_stage_arrays(c);
        if (which == STAGE_BC_CELLS) require_bcs(c);
        StageArgs a = stage_args(c);
        a.only_bc = bc;
        HIPCHK(stage_launch(which, a, 0.0, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_set_gamma(hcfd_ctx* c, double gamma) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        if (!(gamma > 1.0)) fail(HCFD_ERR_ARG, "gamma must exceed 1");
        if (c->bcs_set && gamma != c->gamma) fail(HCFD_ERR_STATE, "gamma differs from the one the BCs were set with");
        c->gamma = gamma;
        c->bound_valid = false;
        drop_graph(c);
    });
}

int hcfd_step(hcfd_ctx* c, int nsteps, double dt, double cfl, double* t_inout) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        run_steps(c, nsteps, dt, cfl);
        c->t_host = device_time(c);  // the kernels advance the clocks in both dt modes
        if (t_inout) *t_inout = c->t_host;
    });
}

int hcfd_step_async(hcfd_ctx* c, int nsteps, double dt, double cfl) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        run_steps(c, nsteps, dt, cfl);
    });
}

int hcfd_step_async_noalloc(hcfd_ctx* c, int nsteps, double dt, double cfl) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        run_steps_noalloc(c, nsteps, dt, cfl);
    });
}

int hcfd_sync(hcfd_ctx* c, double* t) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        c->t_host = device_time(c);
        if (t) *t = c->t_host;
    });
}

int hcfd_output_begin(hcfd_ctx* c, int slot) {
    return guarded([&] {
        if (!c || slot < 0 || slot > 1) fail(HCFD_ERR_ARG, "slot must be 0 or 1");
        if (c->out_pending[slot]) fail(HCFD_ERR_STATE, "output slot still pending: hcfd_output_wait first");
        const size_t n = 4 * (size_t)c->Nc + 1;  // primitives [4][Nc], then the clock
        if (!c->cst) HIPCHK(hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking));
        if (!c->out_host[slot]) {
            // all four resources or none: a failure leaves the slot unallocated
            if (!c->out_dev[slot]) c->out_dev[slot] = dmalloc<double>(n);
            if (!c->out_ready[slot]) HIPCHK(hipEventCreateWithFlags(&c->out_ready[slot], hipEventDisableTiming));
            if (!c->out_copied[slot]) HIPCHK(hipEventCreateWithFlags(&c->out_copied[slot], hipEventDisableTiming));
            void* h = nullptr;
            if (hipHostMalloc(&h, sizeof(double) * n, hipHostMallocDefault) != hipSuccess || !h) {
                (void)hipGetLastError();
                fail(HCFD_ERR_NOMEM, "pinned host allocation of an output slot failed");
            }
            c->out_host[slot] = static_cast<double*>(h);
        }
        double* d = c->out_dev[slot];
        HIPCHK(launch_prims(c->sched.aosoa_width, c->Nc, c->d_old_of_new, c->U[c->cur], c->gamma, d, c->Nc, c->st));
        HIPCHK(hipMemcpyAsync(d + 4 * (size_t)c->Nc, c->tdev, sizeof(double), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipEventRecord(c->out_ready[slot], c->st));
        HIPCHK(hipStreamWaitEvent(c->cst, c->out_ready[slot], 0));
        HIPCHK(hipMemcpyAsync(c->out_host[slot], d, sizeof(double) * n, hipMemcpyDeviceToHost, c->cst));
        HIPCHK(hipEventRecord(c->out_copied[slot], c->cst));
        c->out_pending[slot] = true;
    });
}

int hcfd_output_wait(hcfd_ctx* c, int slot, double* const Q[4], double* t) {
    return guarded([&] {
        if (!c || slot < 0 || slot > 1) fail(HCFD_ERR_ARG, "slot must be 0 or 1");
        if (!c->out_pending[slot]) fail(HCFD_ERR_STATE, "output slot was not begun");
        c->out_pending[slot] = false;
        HIPCHK(hipEventSynchronize(c->out_copied[slot]));
        const double* h = c->out_host[slot];
        if (Q)
            for (int v = 0; v < NV; v++) {
                if (!Q[v]) fail(HCFD_ERR_ARG, "null array");
                std::memcpy(Q[v], h + (size_t)v * c->Nc, sizeof(double) * c->Nc);
            }
        if (t) *t = h[4 * (size_t)c->Nc];
    });
}

int hcfd_get_last_dt(hcfd_ctx* c, double* dt) {
    return guarded([&] {
        if (!c || !dt) fail(HCFD_ERR_ARG, "null");
        HIPCHK(hipMemcpyAsync(dt, c->lastdt, sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_get_fault(hcfd_ctx* c, int* fault) {
    return guarded([&] {
        if (!c || !fault) fail(HCFD_ERR_ARG, "null");
        std::vector<int> f(c->ninst);
        HIPCHK(hipMemcpyAsync(f.data(), c->fault, sizeof(int) * c->ninst, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        *fault = 0;
        for (int x : f) *fault |= x;
    });
}

int hcfd_time_steps(hcfd_ctx* c, int nsteps, double dt, double cfl, float* ms_out) {
    return guarded([&] {
        if (!c || !ms_out || nsteps <= 0) fail(HCFD_ERR_ARG, "bad argument");
        require_bcs(c);
        if (!(cfl > 0)) load_fixed_dt(c, dt);
        else c->fixed_dt_loaded = NAN;
        ensure_bound(c, cfl);
        // One timed region: nsteps eager steps with an event pair around every
        // launch; wall = first to last event.
        std::vector<hipEvent_t> ev(4 * nsteps);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int s = 0; s < nsteps; s++) c->cur = enqueue_step(c, c->cur, cfl, &ev[4 * s]);
        HIPCHK(hipEventSynchronize(ev[4 * nsteps - 1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[4 * nsteps - 1]));
        ms_out[0] = ms;
        double sa = 0, sr = 0, sb = 0;
        for (int s = 0; s < nsteps; s++) {
            float a = 0, r = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, ev[4 * s], ev[4 * s + 1]));
            HIPCHK(hipEventElapsedTime(&r, ev[4 * s + 1], ev[4 * s + 2]));
            HIPCHK(hipEventElapsedTime(&b, ev[4 * s + 2], ev[4 * s + 3]));
            sa += a;
            sr += r;
            sb += b;
        }
        ms_out[1] = (float)(sa / nsteps);
        ms_out[2] = (float)(sb / nsteps);
        ms_out[3] = (cfl > 0) ? (float)(sr / nsteps) : 0.0f;
        for (auto& e : ev) (void)hipEventDestroy(e);
        c->t_host = device_time(c);
    });
}

int hcfd_copy_bandwidth(long long bytes, int reps, double* gbs) {
    return guarded([&] {
        if (bytes < 16 || reps <= 0 || !gbs) fail(HCFD_ERR_ARG, "bad argument");
        const long long n16 = bytes / 16;
        if ((n16 + 255) / 256 > 0x7fffffffLL) fail(HCFD_ERR_ARG, "copy too large");
        void* a = nullptr;
        void* b = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        try {
            HIPCHK(hipMalloc(&a, n16 * 16));
            HIPCHK(hipMalloc(&b, n16 * 16));
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HIPCHK(hipMemsetAsync(a, 0, n16 * 16, st));
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            const int grid = (int)((n16 + 255) / 256);  // one element per lane
            HIPCHK(launch_copy16(n16, a, b, grid, st));  // warm-up
            HIPCHK(hipEventRecord(e0, st));
            for (int r = 0; r < reps; r++) HIPCHK(launch_copy16(n16, a, b, grid, st));
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            *gbs = 2.0 * 16.0 * n16 * reps / (ms * 1e-3) / 1e9;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int hcfd_set_schedule(hcfd_ctx* c, const hcfd_schedule* s) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        hcfd_schedule n = s ? *s : hcfd_schedule{256, 1, 1, 0, 1, c->fused_default ? 1 : 0};
        check_schedule(c, n);
        change_width(c, n.aosoa_width);
        c->sched = n;
        drop_graph(c);
    });
}

int hcfd_set_integrator(hcfd_ctx* c, int stages) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        if (stages < 1 || stages > 3) fail(HCFD_ERR_ARG, "stages must be 1, 2 or 3");
        if (stages > 1 && !c->U[2]) {
            const size_t n = 4 * (size_t)c->NcPad;  // as U[0], U[1]
            c->U[2] = dmalloc<double>(n);
            HIPCHK(hipMemsetAsync(c->U[2], 0, sizeof(double) * n, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        }
        c->stages = stages;
        drop_graph(c);
    });
}

int hcfd_get_integrator(hcfd_ctx* c, int* stages) {
    return guarded([&] {
        if (!c || !stages) fail(HCFD_ERR_ARG, "null");
        *stages = c->stages;
    });
}

int hcfd_get_schedule(hcfd_ctx* c, hcfd_schedule* s) {
    return guarded([&] {
        if (!c || !s) fail(HCFD_ERR_ARG, "null");
        *s = c->sched;
    });
}

int hcfd_autotune(hcfd_ctx* c, int steps, double cfl, const char* cache_path, hcfd_schedule* best) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        require_bcs(c);
        if (steps <= 0) steps = 10;
        // the key holds what decides the candidates' speed: sizes, precision, cell
        // order (flags other than graph replay), slot bucket, integrator stages
        char key[200];
        snprintf(key, sizeof key, "gfx950 Nc=%d Nf=%d Nint=%d ninst=%d prec=%d flags=%d deg=%d stages=%d fused_ok=%d",
                 c->Nc, c->Nf, c->Nint, c->ninst, c->precision, c->flags & ~HCFD_FLAG_GRAPH, c->deg_bucket, c->stages,
                 c->fused_ok ? 1 : 0);
        // cache lookup
        if (cache_path) {
            std::ifstream in(cache_path);
            std::string line;
            while (std::getline(in, line)) {
                auto bar = line.find('|');
                if (bar == std::string::npos || line.substr(0, bar) != key) continue;
                hcfd_schedule s{};
                if (sscanf(line.c_str() + bar + 1, "%d %d %d %d %d %d", &s.block, &s.lanes_per_cell,
                           &s.aosoa_width, &s.persistent, &s.xcd_remap, &s.fused) == 6) {
                    try {
                        check_schedule(c, s);
                    } catch (const Error&) {
                        continue;  // an entry this context cannot run: tune instead
                    }
                    change_width(c, s.aosoa_width);
                    c->sched = s;
                    drop_graph(c);
                    if (best) *best = s;
                    return;
                }
            }
        }
        const size_t P = c->NcPad;
        if (!c->backup) c->backup = dmalloc<double>(4 * P);
        // clocks, last dt and fault flags are restored after the probes
        std::vector<double> t0(c->ninst), l0(c->ninst);
        std::vector<int> f0(c->ninst);
        HIPCHK(hipMemcpyAsync(t0.data(), c->tdev, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(l0.data(), c->lastdt, sizeof(double) * c->ninst, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(f0.data(), c->fault, sizeof(int) * c->ninst, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        const double dt_probe = 1e-12;  // only used when cfl <= 0: a negligible step keeps the state valid
        std::vector<hcfd_schedule> cands;
        for (int w : {1, 8, 64})
            for (int b : {64, 128, 256})
                for (int lpc : {1, 4})
                    for (int pers : {0, 8})
                        for (int x : {0, 1}) {
                            cands.push_back(hcfd_schedule{b, lpc, w, pers, x, 0});
                            if (c->fused_ok && b == 256 && lpc == 1) cands.push_back(hcfd_schedule{b, lpc, w, pers, x, 1});
                        }
        const double t_save = c->t_host;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        float best_ms = 1e30f;
        hcfd_schedule bs = c->sched;
        for (const auto& s : cands) {
            change_width(c, s.aosoa_width);
            c->sched = s;
            c->bound_valid = false;
            drop_graph(c);
            HIPCHK(hipMemcpyAsync(c->backup, c->U[c->cur], sizeof(double) * 4 * P, hipMemcpyDeviceToDevice, c->st));
            const int cur0 = c->cur;
            run_steps(c, 2, dt_probe, cfl);  // warm-up
            HIPCHK(hipEventRecord(e0, c->st));
            run_steps(c, steps, dt_probe, cfl);
            HIPCHK(hipEventRecord(e1, c->st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            c->cur = cur0;
            HIPCHK(hipMemcpyAsync(c->U[c->cur], c->backup, sizeof(double) * 4 * P, hipMemcpyDeviceToDevice, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            if (ms < best_ms) {
                best_ms = ms;
                bs = s;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        change_width(c, bs.aosoa_width);
        c->sched = bs;
        drop_graph(c);
        c->t_host = t_save;
        // dt/time bookkeeping of the probes is discarded
        HIPCHK(hipMemcpyAsync(c->tdev, t0.data(), sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->lastdt, l0.data(), sizeof(double) * c->ninst, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->fault, f0.data(), sizeof(int) * c->ninst, hipMemcpyHostToDevice, c->st));
        c->fixed_dt_loaded = NAN;
        c->bound_valid = false;
        HIPCHK(hipStreamSynchronize(c->st));
        if (cache_path) {
            std::ofstream out(cache_path, std::ios::app);
            out << key << "|" << bs.block << " " << bs.lanes_per_cell << " " << bs.aosoa_width << " "
                << bs.persistent << " " << bs.xcd_remap << " " << bs.fused << "  # " << best_ms / steps << " ms/step\n";
        }
        if (best) *best = bs;
    });
}

int hcfd_batch_get_state(hcfd_ctx* c, int inst, double* const U[4]) {
    return guarded([&] {
        if (!c || !U || inst < 0 || inst >= c->ninst) fail(HCFD_ERR_ARG, "bad argument");
        HIPCHK(launch_export(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->U[c->cur], c->staging, c->Nc, c->st));
        const int o = c->inst_cell_off[inst], n = c->inst_cell_off[inst + 1] - o;
        for (int v = 0; v < NV; v++)
            HIPCHK(hipMemcpyAsync(U[v], c->staging + (size_t)v * c->Nc + o, sizeof(double) * n, hipMemcpyDeviceToHost,
                                  c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_batch_get_prims(hcfd_ctx* c, int inst, double* const Q[4]) {
    return guarded([&] {
        if (!c || !Q || inst < 0 || inst >= c->ninst) fail(HCFD_ERR_ARG, "bad argument");
        HIPCHK(launch_prims(c->sched.aosoa_width, c->Nc, c->d_old_of_new, c->U[c->cur], c->gamma, c->staging, c->Nc,
                            c->st));
        const int o = c->inst_cell_off[inst], n = c->inst_cell_off[inst + 1] - o;
        for (int v = 0; v < NV; v++)
            HIPCHK(hipMemcpyAsync(Q[v], c->staging + (size_t)v * c->Nc + o, sizeof(double) * n, hipMemcpyDeviceToHost,
                                  c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_batch_set_state(hcfd_ctx* c, int inst, const double* const U[4]) {
    return guarded([&] {
        if (!c || !U || inst < 0 || inst >= c->ninst) fail(HCFD_ERR_ARG, "bad argument");
        HIPCHK(launch_export(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->U[c->cur], c->staging, c->Nc, c->st));
        const int o = c->inst_cell_off[inst], n = c->inst_cell_off[inst + 1] - o;
        for (int v = 0; v < NV; v++)
            HIPCHK(hipMemcpyAsync(c->staging + (size_t)v * c->Nc + o, U[v], sizeof(double) * n, hipMemcpyHostToDevice,
                                  c->st));
        HIPCHK(launch_import(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->staging, c->Nc, c->U[c->cur], c->st));
        c->bound_valid = false;
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int hcfd_batch_get_status(hcfd_ctx* c, double* t, double* last_dt, int* fault) {
    return guarded([&] {
        if (!c) fail(HCFD_ERR_ARG, "null");
        const int n = c->ninst;
        if (t) {
            HIPCHK(hipMemcpyAsync(t, c->tdev, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
        }
        if (last_dt) HIPCHK(hipMemcpyAsync(last_dt, c->lastdt, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
        if (fault) HIPCHK(hipMemcpyAsync(fault, c->fault, sizeof(int) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

// ---- checkpoint / resume ------------------------------------------------------
// File: "HCFDCKP1", int32 version (1), int32 ninst, int64 Nc (all instances,
// ghosts included), int32 precision, int32 stages, int32 Nc of each instance,
// float64 clock and last dt of each instance, int32 fault flag of each, then U
// as four float64 arrays of Nc values in the caller's cell order (instances
// one after another), then a 64-bit FNV-1a hash of those U bytes.  Host byte
// order (little-endian on the platforms this library builds for).
namespace {

constexpr char CKP_MAGIC[8] = {'H', 'C', 'F', 'D', 'C', 'K', 'P', '1'};

struct Fnv1a {
    unsigned long long h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    }
};

struct File {
    std::FILE* f = nullptr;
    std::string path;
    File(const char* p, const char* mode) : f(std::fopen(p, mode)), path(p) {
        if (!f) fail(HCFD_ERR_IO, std::string("cannot open ") + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) fail(HCFD_ERR_IO, path + ": write failed");
    }
    void read(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) fail(HCFD_ERR_IO, path + ": truncated checkpoint");
    }
    void close() {
        const int r = std::fclose(f);
        f = nullptr;
        if (r != 0) fail(HCFD_ERR_IO, path + ": close failed");
    }
};

}  // namespace

int hcfd_checkpoint_save(hcfd_ctx* c, const char* path) {
    return guarded([&] {
        if (!c || !path) fail(HCFD_ERR_ARG, "null");
        const int n = c->ninst;
        std::vector<double> t(n), dt(n);
        std::vector<int> fault(n), sizes(n);
        HIPCHK(hipMemcpyAsync(t.data(), c->tdev, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(dt.data(), c->lastdt, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(fault.data(), c->fault, sizeof(int) * n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(launch_export(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->U[c->cur], c->staging, c->Nc, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int k = 0; k < n; k++) sizes[k] = c->inst_cell_off[k + 1] - c->inst_cell_off[k];
        File f(path, "wb");
        const int head[2] = {1, n};
        const long long Nc = c->Nc;
        const int pre[2] = {c->precision, c->stages};
        f.write(CKP_MAGIC, sizeof CKP_MAGIC);
        f.write(head, sizeof head);
        f.write(&Nc, sizeof Nc);
        f.write(pre, sizeof pre);
        f.write(sizes.data(), sizeof(int) * n);
        f.write(t.data(), sizeof(double) * n);
        f.write(dt.data(), sizeof(double) * n);
        f.write(fault.data(), sizeof(int) * n);
        // one variable at a time through a host buffer of Nc values
        std::vector<double> h((size_t)c->Nc);
        Fnv1a hash;
        for (int v = 0; v < NV; v++) {
            HIPCHK(hipMemcpy(h.data(), c->staging + (size_t)v * c->Nc, sizeof(double) * c->Nc,
                             hipMemcpyDeviceToHost));
            hash.add(h.data(), sizeof(double) * h.size());
            f.write(h.data(), sizeof(double) * h.size());
        }
        f.write(&hash.h, sizeof hash.h);
        f.close();
    });
}

int hcfd_checkpoint_load(hcfd_ctx* c, const char* path) {
    return guarded([&] {
        if (!c || !path) fail(HCFD_ERR_ARG, "null");
        HIPCHK(hipStreamSynchronize(c->st));  // steps still in flight read the state and the staging buffer
        const int n = c->ninst;
        File f(path, "rb");
        char magic[8];
        int head[2], pre[2];
        long long Nc = 0;
        f.read(magic, sizeof magic);
        if (std::memcmp(magic, CKP_MAGIC, sizeof magic) != 0) fail(HCFD_ERR_IO, f.path + ": not a checkpoint");
        f.read(head, sizeof head);
        if (head[0] != 1) fail(HCFD_ERR_IO, f.path + ": checkpoint version " + std::to_string(head[0]));
        f.read(&Nc, sizeof Nc);
        f.read(pre, sizeof pre);
        if (head[1] != n || Nc != c->Nc)
            fail(HCFD_ERR_ARG, f.path + ": checkpoint of " + std::to_string(head[1]) + " instance(s), " +
                                   std::to_string(Nc) + " cells; this context has " + std::to_string(n) + ", " +
                                   std::to_string(c->Nc));
        std::vector<int> sizes(n), fault(n);
        std::vector<double> t(n), dt(n);
        f.read(sizes.data(), sizeof(int) * n);
        for (int k = 0; k < n; k++)
            if (sizes[k] != c->inst_cell_off[k + 1] - c->inst_cell_off[k])
                fail(HCFD_ERR_ARG, f.path + ": instance " + std::to_string(k) + " has " + std::to_string(sizes[k]) +
                                       " cells in the checkpoint, " +
                                       std::to_string(c->inst_cell_off[k + 1] - c->inst_cell_off[k]) + " here");
        f.read(t.data(), sizeof(double) * n);
        f.read(dt.data(), sizeof(double) * n);
        f.read(fault.data(), sizeof(int) * n);
        std::vector<double> h((size_t)c->Nc);
        Fnv1a hash;
        for (int v = 0; v < NV; v++) {
            f.read(h.data(), sizeof(double) * h.size());
            hash.add(h.data(), sizeof(double) * h.size());
            HIPCHK(hipMemcpy(c->staging + (size_t)v * c->Nc, h.data(), sizeof(double) * c->Nc,
                             hipMemcpyHostToDevice));
        }
        unsigned long long stored = 0;
        f.read(&stored, sizeof stored);
        if (stored != hash.h) fail(HCFD_ERR_IO, f.path + ": checkpoint state does not match its hash");
        // nothing on the device has changed yet: the state, clocks and flags go in together
        HIPCHK(launch_import(c->sched.aosoa_width, NV, c->Nc, c->d_old_of_new, c->staging, c->Nc, c->U[c->cur], c->st));
        HIPCHK(hipMemcpyAsync(c->tdev, t.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->lastdt, dt.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->fault, fault.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->t_host = t[0];
        c->bound_valid = false;  // the next CFL step takes its dt from the loaded state
    });
}

int hcfd_batch_instance_size(hcfd_ctx* c, int inst, int* Nc, int* Nf) {
    return guarded([&] {
        if (!c || inst < 0 || inst >= c->ninst) fail(HCFD_ERR_ARG, "bad argument");
        if (Nc) *Nc = c->inst_cell_off[inst + 1] - c->inst_cell_off[inst];
        if (Nf) *Nf = c->inst_face_off[inst + 1] - c->inst_face_off[inst];
    });
}

}  // extern "C"
