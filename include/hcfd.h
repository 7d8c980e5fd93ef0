/*
 * hcfd.h -- C ABI of the MI355X (gfx950) explicit unstructured finite-volume
 * Euler residual library (libhcfd.so).
 *
 * The reference selects a layout by namespace (Base, SoA, VEC, AoSoA; see
 * src/Test.cpp:64-74) and calls the layout's Domain / GaussGreen / Limiter /
 * Flux / BoundaryCondition / Euler methods.  The C++ classes of namespace
 * MI355X (cfd_optimisations_amd/host/) keep those shapes and call the entry
 * points below; each declaration names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status (HCFD_OK == 0); hcfd_last_error()
 *     returns the text of the calling thread's most recent failure.
 *   - Host arrays are caller-owned and only touched during the call.
 *     Per-variable arrays are passed as `double* const X[4]` (one pointer per
 *     conserved/primitive variable, Nc entries each, in the caller's original
 *     cell order), the reference's `double* U[NVARS]` convention.
 *     Gradients are `X[v][d*Nc + i]` (includes/Base/GradientSchemes.hpp).
 *   - A context owns its device buffers and one HIP stream; it may be used by
 *     one host thread at a time.  Cells are renumbered internally (reverse
 *     Cuthill-McKee, or a Hilbert curve with HCFD_FLAG_ORDER_SFC); every
 *     getter/setter speaks the caller's order.
 *   - No C++ exceptions cross this boundary.
 */
#ifndef HCFD_H
#define HCFD_H

#ifdef __cplusplus
extern "C" {
#endif

#define HCFD_NVARS 4
#define HCFD_NDIMS 2

enum hcfd_status {
    HCFD_OK = 0,
    HCFD_ERR_ARG = 1,     /* bad argument / size out of range */
    HCFD_ERR_MESH = 2,    /* mesh violates an invariant (e.g. ghost not on the right) */
    HCFD_ERR_HIP = 3,     /* HIP runtime error */
    HCFD_ERR_STATE = 4,   /* call out of order (e.g. step before set_bcs) */
    HCFD_ERR_IO = 5,      /* file could not be read / written / truncated */
    HCFD_ERR_NOMEM = 6,   /* host or device allocation failed */
    HCFD_ERR_NUMERIC = 7  /* non-physical state detected (rho <= 0, p <= 0 or NaN) */
};

/* Arithmetic of the residual. */
enum hcfd_precision {
    HCFD_FP64 = 0,  /* plain FP64 path: the drop-in for Base */
    HCFD_MIXED = 1  /* FP32 differences from the FP64 state, gradient, limiter and flux;
                       residual summed in FP64; FP64 state, update and dt */
};

/* Same values and order as Base::BCtype (includes/Base/BoundaryConditions.hpp:6-13). */
enum hcfd_bctype {
    HCFD_SYMMETRY = 0,
    HCFD_WALL = 1,
    HCFD_SUPINLET = 2,
    HCFD_SUBINLET = 3,
    HCFD_SUPOUTLET = 4,
    HCFD_SUBOUTLET = 5
};

/* Creation flags. */
enum hcfd_flags {
    HCFD_FLAG_NONE = 0,
    HCFD_FLAG_EXACT = 1,       /* accepted for compatibility: the FP64 path is always bit-identical to Base */
    HCFD_FLAG_NO_RENUMBER = 2, /* keep the caller's cell order (no RCM); faces still grouped by cell */
    HCFD_FLAG_GRAPH = 4,       /* replay multi-step runs from a captured hipGraph */
    HCFD_FLAG_ORDER_SFC = 8    /* order interior cells along a Hilbert curve of their centroids
                                  (compact blocks) instead of reverse Cuthill-McKee */
};

/* Raw mesh arrays, exactly the members the reference's Mesh reads from a
 * .hmsh file (includes/Mesh.hpp:4-22, src/Mesh.cpp:11-45). */
typedef struct hcfd_mesh_view {
    int dim, Nn, Ncol, NBCcol, Nc, Nf;
    const int* coltag;   /* [Ncol] */
    const int* coloff;   /* [Ncol+1] */
    const double* nodes; /* [Nn*2] */
    const int* fn_con;   /* [2*Nf] */
    const int* fc_con;   /* [2*Nf] */
    const int* cn_off;   /* [Nc+1] */
    const int* cn_con;   /* [cn_off[Nc]] */
} hcfd_mesh_view;

/* One case of a batch: boundary conditions by BC ordinal, gas constant and a
 * uniform primitive initial state (Euler::gather_BCs/initialise). */
typedef struct hcfd_case {
    const int* bctype;   /* [NBCcol] hcfd_bctype per BC ordinal (colour Ncol-NBCcol+bc) */
    const double* qbc;   /* [NBCcol*4] primitive BC states (rho, u, v, p) */
    double qic[4];       /* uniform primitive initial state */
} hcfd_case;

/* Launch schedule of the two fused passes; filled by the autotuner. */
typedef struct hcfd_schedule {
    int block;           /* threads per block: 64, 128, 256 */
    int lanes_per_cell;  /* pass A lanes per cell: 1, or 4 (one per conserved variable; 16 cells per wavefront) */
    int aosoa_width;     /* cell-state tile width: 1 (AoS), 8, 64 */
    int persistent;      /* 0: one tile per block; >0: blocks per CU of a persistent grid */
    int xcd_remap;       /* 1: map block ids so each XCD walks a contiguous cell range */
    int fused;           /* 1: one kernel per step over tiles of 256 cells that recompute a one-cell
                            halo ring (face states stay in LDS); needs block 256, lanes_per_cell 1
                            and hcfd_info.fused_available */
} hcfd_schedule;

typedef struct hcfd_info {
    int Nc, Nf, Nint, Nghost, ninst, maxdeg;
    int precision, flags;
    int bandwidth_in;    /* max |CL-CR| over interior faces, caller's order */
    int bandwidth_out;   /* same after renumbering */
    int ncolours;        /* face colours of the conflict-free colouring; 0 until
                            hcfd_get_face_colours has built it (on its first call) */
    long long device_bytes;
    hcfd_schedule sched;
    int fused_available;  /* the halo tiles of this mesh fit in LDS and their local-index tables
                             were built (see hcfd_schedule.fused); the
                             single-kernel step is the default schedule when, in addition,
                             at most 5% of the tiles (192 cells in FP64, 256 in FP32-mixed)
                             had to be split to fit */
    int fused_lds_bytes;  /* LDS per block of the single-kernel step */
    int halo_max_cells;   /* most cells one tile stages: its own (at most 256) + halo ring + their neighbours */
} hcfd_info;

typedef struct hcfd_ctx hcfd_ctx;

const char* hcfd_last_error(void);
int hcfd_version(void);
int hcfd_device_count(int* n);

/* ---- context (Domain) --------------------------------------------------- */

/* Replaces Base::Domain::Domain(Mesh&) (src/Base/Domain.cpp:4-24) and
 * Mesh::compute_cf_con (src/Mesh.cpp:88-117): uploads the mesh, builds the
 * cell->face CSR, the RCM (or Hilbert) renumbering and the face/cell
 * geometry on the device, and applies Domain::zero_bc_iv
 * (src/Base/Domain.cpp:148-154) as the Euler constructor does. */
int hcfd_create(const hcfd_mesh_view* mesh, int precision, int flags, hcfd_ctx** out);
int hcfd_destroy(hcfd_ctx* ctx);
int hcfd_get_info(hcfd_ctx* ctx, hcfd_info* out);

/* Geometry in the caller's order, as Base::Domain holds it
 * (includes/Base/Domain.hpp:13-25): rf [Nf][2][2], nf/xf [Nf][2], Af/If [Nf],
 * xc [Nc][2], iv [Nc] (ghost iv = 0 after zero_bc_iv).  Any pointer may be NULL. */
int hcfd_get_geometry(hcfd_ctx* ctx, double* xf, double* xc, double* rf, double* nf, double* Af, double* If,
                      double* iv);
/* cell->face CSR in the caller's numbering (Mesh::cf_off/cf_con, src/Mesh.cpp:88-117). */
int hcfd_get_cf(hcfd_ctx* ctx, int* cf_off, int* cf_con);
/* new_of_old[i] = internal index of caller cell i. */
int hcfd_get_permutation(hcfd_ctx* ctx, int* new_of_old);
/* Conflict-free face colouring (no two faces of one colour share a cell);
 * colour_of_face in the caller's face order.  Replaces the empty
 * Mesh::colour_faces (src/Mesh.cpp:119-121).  Built on the device on the first
 * call (Jones-Plassmann rounds), not at creation: no step kernel needs it. */
int hcfd_get_face_colours(hcfd_ctx* ctx, int* colour_of_face, int* ncolours);

/* Boundary conditions by BC ordinal: Euler::gather_BCs (src/Base/Euler.cpp:42-60)
 * with BoundaryCondition::init (src/Base/BoundaryConditions/Base.cpp:8-23). */
int hcfd_set_bcs(hcfd_ctx* ctx, int nbc, const int* bctype, const double* qbc, double gamma);

/* ---- state ---------------------------------------------------------------- */

/* Euler::initialise (src/Base/Euler.cpp:62-77): uniform state then BC ghosts. */
int hcfd_initialise(hcfd_ctx* ctx, const double qic[4]);
/* Per-cell primitive initial state (same arithmetic), then BC ghosts. */
int hcfd_initialise_prims(hcfd_ctx* ctx, const double* const Q[4]);
int hcfd_set_state(hcfd_ctx* ctx, const double* const U[4]);
int hcfd_get_state(hcfd_ctx* ctx, double* const U[4]);
/* Euler::compute_prims (src/Base/Euler.cpp:115-136), computed on the device. */
int hcfd_get_prims(hcfd_ctx* ctx, double* const Q[4]);

/* ---- per-stage entry points (drop-in parity with the Base classes) -------- */
/* These act on device-resident stage arrays; the hcfd_stage_get/set calls
 * move them.  which: 0 = U, 1 = G ([4][2*Nc]), 2 = phi, 3 = R.  They are the
 * FP64 Base methods: an HCFD_MIXED context returns HCFD_ERR_STATE. */
enum hcfd_array { HCFD_ARR_U = 0, HCFD_ARR_G = 1, HCFD_ARR_PHI = 2, HCFD_ARR_R = 3 };
int hcfd_stage_set(hcfd_ctx* ctx, int which, const double* const X[4]);
int hcfd_stage_get(hcfd_ctx* ctx, int which, double* const X[4]);
/* GaussGreen::compute_gradient (src/Base/GaussGreen.cpp:14-40): G = grad(U). */
int hcfd_stage_gradient(hcfd_ctx* ctx);
/* Limiter::limit_gradient (src/Base/Limiter.cpp:19-77): phi = lim(U, G). */
int hcfd_stage_limiter(hcfd_ctx* ctx);
/* BoundaryCondition::update_grads / update_limiter
 * (src/Base/BoundaryConditions/Base.cpp:29-53), all BCs. */
int hcfd_stage_bc_grads(hcfd_ctx* ctx);
int hcfd_stage_bc_limiter(hcfd_ctx* ctx);
/* Flux::ComputeFlux (src/Base/Flux.cpp:9-50): R = (R + sum of face fluxes) * iv. */
int hcfd_stage_flux(hcfd_ctx* ctx);
/* Euler::add_residuals + clear (src/Base/Euler.cpp:107-113, :93-95): U += dt R; R = 0. */
int hcfd_stage_update(hcfd_ctx* ctx, double dt);
/* BoundaryCondition::update_cells for every BC (Euler::update_BC_cells,
 * src/Base/Euler.cpp:138-142). */
int hcfd_stage_bc_cells(hcfd_ctx* ctx);
/* One BoundaryCondition object's update_cells / update_grads / update_limiter
 * (src/Base/BoundaryConditions/{Symmetry,Wall,Inlet,Outlet}.cpp) on its own faces: BC ordinal bc (-1 = all).
 * what: 0 = cells, 1 = gradients, 2 = limiter. */
int hcfd_stage_bc(hcfd_ctx* ctx, int what, int bc);
/* Ratio of specific heats used by the stage flux (Flux::Flux(nvars, gma, ...),
 * includes/Base/Flux.hpp:9) before any BCs are set. */
int hcfd_set_gamma(hcfd_ctx* ctx, double gamma);

/* ---- fused stepping (the hot path) --------------------------------------- */

/* nsteps steps of Euler::run's loop body (src/Base/Euler.cpp:85-97) as two
 * fused passes (or one single-kernel launch) per step and stage.  cfl > 0: dt = cfl * min_i 1/(iv_i sum_f A_f(|u.n|+c))
 * recomputed every step (no reference counterpart); otherwise fixed dt.
 * t_inout (may be NULL) is advanced by the sum of the steps' dt. */
int hcfd_step(hcfd_ctx* ctx, int nsteps, double dt, double cfl, double* t_inout);
/* hcfd_step without the wait: enqueues the steps on the context's stream and
 * returns; hcfd_sync waits for them and gives the simulated time (t may be
 * NULL).  Lets the host write output while the device steps on. */
int hcfd_step_async(hcfd_ctx* ctx, int nsteps, double dt, double cfl);
int hcfd_sync(hcfd_ctx* ctx, double* t);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): hcfd_step_async with nothing allocated, on the device or the
 * host.  The steps' CFL scratch (per-tile partial maxima and per-instance
 * atomics, which hcfd_step_async keeps in buffers the context allocated) lives
 * in one static 16 MiB device array per GPU, held until the process ends; a
 * context whose schedule needs more (2 x tiles + instances + 1 doubles: about a
 * million tiles, some 200M cells in FP64's 192-cell tiles and 260M in
 * FP32-mixed's 256-cell ones) gets HCFD_ERR_ARG.  Launches are eager (no graph is
 * instantiated); a fixed dt is stored by a kernel.  Results equal
 * hcfd_step_async's bit for bit.  Calls that overlap on one GPU share the array
 * and race: run them on one stream, or one after another. */
int hcfd_step_async_noalloc(hcfd_ctx* ctx, int nsteps, double dt, double cfl);
/* Output while stepping (the reference computes primitives and writes them
 * with the solver stopped, Euler::run, src/Base/Euler.cpp:99-103):
 * hcfd_output_begin enqueues Euler::compute_prims (src/Base/Euler.cpp:115-136)
 * of the current state into device slot `slot` (0 or 1), then, on a second
 * stream, its copy to pinned host memory, and returns at once; steps enqueued
 * afterwards overlap the copy.  hcfd_output_wait blocks until the slot's copy
 * has landed and gives the primitives (caller order, all cells) and the
 * simulated time of the snapshot (Q or t may be NULL).  Each begin needs its
 * wait before the slot is begun again. */
int hcfd_output_begin(hcfd_ctx* ctx, int slot);
int hcfd_output_wait(hcfd_ctx* ctx, int slot, double* const Q[4], double* t);
/* Checkpoint / resume (no counterpart in the reference, which can restart only
 * from the primitives of a .hsln file, src/Base/Euler.cpp:115-136, so not bit for
 * bit).  hcfd_checkpoint_save writes the conserved state of every cell (caller
 * order, ghosts included) and each instance's clock, last dt and fault flag, in
 * FP64, with a hash of the state; hcfd_checkpoint_load restores them into a
 * context built from the same mesh(es) with its BCs set (any precision,
 * integrator or schedule), after which stepping continues exactly as the saved
 * run would have.  A checkpoint of other meshes, a truncated file or a hash
 * mismatch is refused with HCFD_ERR_ARG / HCFD_ERR_IO and leaves the context
 * unchanged. */
int hcfd_checkpoint_save(hcfd_ctx* ctx, const char* path);
int hcfd_checkpoint_load(hcfd_ctx* ctx, const char* path);
/* dt of the most recent step. */
int hcfd_get_last_dt(hcfd_ctx* ctx, double* dt);
/* 1 if any step since the last call produced rho <= 0, p <= 0 or NaN. */
int hcfd_get_fault(hcfd_ctx* ctx, int* fault);

/* Time integrator of hcfd_step / hcfd_time_steps: stages = 1 forward Euler
 * (Euler::run, the reference's only integrator, includes/Base/TimeIntegrators.hpp:11-45),
 * 2 SSP-RK2, 3 SSP-RK3 in Shu-Osher form (no reference counterpart):
 *   RK2: U1 = E(U^n); U^{n+1} = 1/2 U^n + 1/2 E(U1)
 *   RK3: U1 = E(U^n); U2 = 3/4 U^n + 1/4 E(U1); U^{n+1} = 1/3 U^n + 2/3 E(U2)
 * with E(W) = W + dt R(W) the forward-Euler stage (ghost cells set from each
 * stage's interior) and dt the fixed dt or the CFL step of U^n.  Each stage is
 * one more pass over the mesh; choosing 2 or 3 allocates a third state buffer. */
int hcfd_set_integrator(hcfd_ctx* ctx, int stages);
int hcfd_get_integrator(hcfd_ctx* ctx, int* stages);

/* Runs nsteps steps on the context's stream with a HIP event pair around
 * every launch.  ms_out[0] = first to last event (wall of the nsteps steps);
 * ms_out[1] = mean pass A (gradient + limiter (+ CFL bound)), or mean of the
 * single-kernel step when the schedule is fused; ms_out[2] = mean pass B
 * (flux + update + ghost BCs), 0 when fused; ms_out[3] = mean CFL reduction
 * (0 for fixed dt).  With a multi-stage integrator ms_out[1] and ms_out[3] are
 * those of the first stage and ms_out[2] the rest of the step. */
int hcfd_time_steps(hcfd_ctx* ctx, int nsteps, double dt, double cfl, float* ms_out);

/* Device copy ceiling: a plain 16-byte-per-lane copy kernel over `bytes` bytes
 * (read + write), run reps times on a private stream; *gbs = 2*bytes/time.
 * Also the known-size kernel for calibrating HBM byte counters. */
int hcfd_copy_bandwidth(long long bytes, int reps, double* gbs);

/* Sets the launch schedule; NULL restores the default. */
int hcfd_set_schedule(hcfd_ctx* ctx, const hcfd_schedule* s);
int hcfd_get_schedule(hcfd_ctx* ctx, hcfd_schedule* s);
/* Project autotuner: times candidate schedules for `steps` steps each on the
 * current state (state, clock and last dt restored afterwards), keeps the
 * fastest and, when cache_path is non-NULL, records it there keyed by mesh
 * size, batch size, precision, flags, cell degree, integrator stages and
 * whether the single kernel fits.  A cached entry is used without timing; one
 * this context cannot run is ignored and the context is tuned afresh. */
int hcfd_autotune(hcfd_ctx* ctx, int steps, double cfl, const char* cache_path, hcfd_schedule* best);

/* ---- batches of independent instances (one launch per pass for all) ------- */

/* Concatenates n meshes (each with its own case) into one context: cells and
 * faces of all instances are renumbered into one address space, each pass is
 * one launch over all instances, dt is reduced per instance.  The batch is an
 * ordinary context; per-instance getters take the instance index. */
int hcfd_batch_create(int n, const hcfd_mesh_view* meshes, const hcfd_case* cases, double gamma, int precision,
                      int flags, hcfd_ctx** out);
int hcfd_batch_get_state(hcfd_ctx* ctx, int inst, double* const U[4]);
int hcfd_batch_get_prims(hcfd_ctx* ctx, int inst, double* const Q[4]);
int hcfd_batch_set_state(hcfd_ctx* ctx, int inst, const double* const U[4]);
/* Per-instance time, last dt and fault flag; arrays of length ninst (any may be NULL). */
int hcfd_batch_get_status(hcfd_ctx* ctx, double* t, double* last_dt, int* fault);
int hcfd_batch_instance_size(hcfd_ctx* ctx, int inst, int* Nc, int* Nf);

/* ---- the reference's NDIMS == 3 branches ---------------------------------- */
/* Five variables (rho, rho u, rho v, rho w, E), AoS arrays of n entries; FP64 with
 * the reference's expression order (bit-identical to it).  The .hmsh format and
 * the reference's Domain are two-dimensional, so there is no 3-D mesh to step:
 * these evaluate the 3-D arithmetic on the GPU for given states.
 * Flux::Rusanov with the 3-D Flux::EulerFlux (src/Base/Flux.cpp:52-81, 89-104):
 * F[f] = Rusanov(UL[f], UR[f], nrm[f]) for face-state pairs UL, UR [n][5] and
 * unit normals nrm [n][3]. */
int hcfd_rusanov3d(int n, double gamma, const double* UL, const double* UR, const double* nrm, double* F);
/* Euler::compute_prims (src/Base/Euler.cpp:115-136): U [n][5] -> (rho, u, v, w, p). */
int hcfd_prims3d(int n, double gamma, const double* U, double* Q);
/* Euler::initialise's state conversion (src/Base/Euler.cpp:62-68): (rho, u, v, w, p) -> U. */
int hcfd_cons3d(int n, double gamma, const double* Q, double* U);

/* ---- host mesh services (CPU; used by the C++ host classes and tools) ----- */

typedef struct hcfd_mesh hcfd_mesh;
/* Mesh::Mesh(const char*) (src/Mesh.cpp:7-52) with I/O and invariant checks. */
int hcfd_mesh_load(const char* path, hcfd_mesh** out);
int hcfd_mesh_save(const hcfd_mesh_view* mesh, const char* path);
int hcfd_mesh_view_of(const hcfd_mesh* m, hcfd_mesh_view* out);
int hcfd_mesh_free(hcfd_mesh* m);
/* Names of the physical groups (Mesh::parse_pnames, src/Mesh.cpp:70-86): the
 * k-th NUL-terminated string, or NULL. */
const char* hcfd_mesh_pname(const hcfd_mesh* m, int k);

/* Mesh generators (seeded SplitMix64; deterministic across platforms).
 * Riemann: [0,1]^2, nx*ny quads each split in 2 triangles, interior nodes
 * jittered by +-jitter*h, then cells shuffled when shuffle != 0.  Four BC
 * colours (bottom, right, top, left).
 * Ramp: [0,L]x[0,H] channel whose lower wall ramps at angle_deg from x_ramp,
 * mapped structured grid of nx*ny quads split in 2 triangles, jittered; BC
 * colours (lower wall, outlet, upper wall, inlet). */
int hcfd_meshgen_riemann(int nx, int ny, double jitter, unsigned long long seed, int shuffle, hcfd_mesh** out);
int hcfd_meshgen_ramp(int nx, int ny, double length, double height, double x_ramp, double angle_deg,
                      double jitter, unsigned long long seed, int shuffle, hcfd_mesh** out);

/* Writer::write_soln (src/Writer.cpp:22-39): <fname>_<runcount>.hsln with the
 * primitive variables of all Nc cells under the Writer's variable names
 * (src/Writer.cpp:13, 31); varnames NULL gives Density, Velocity_X, Velocity_Y,
 * Pressure. */
int hcfd_write_hsln(const char* path, int Nc, const double* const Q[4], const char* const varnames[4]);
/* The .hsln reader of interface/Export.py:14-32 (read_soln), for restarts:
 * checks nvars == 4 and the file size, sets *Nc and, when Q is non-NULL, fills
 * Q[v][0..Nc).  A restart is hcfd_initialise_prims with the arrays read. */
int hcfd_read_hsln(const char* path, int* Nc, double* const Q[4]);
/* VTU export without the vtk package, the output of interface/Export.py:34-81
 * (write_vtk): all cells (ghost line cells included) with Density, Pressure
 * and Velocity (3 components) as cell data; XML UnstructuredGrid with raw
 * appended binary blocks. */
int hcfd_write_vtu(const char* path, const hcfd_mesh_view* mesh, const double* const Q[4]);

/* gmsh v4 ASCII .msh -> mesh with the reference importer's conventions
 * (interface/Mesh.py:21-127, 159-378): ghost line cells on the right of BC
 * faces, faces grouped by physical tag, RCM cell order. */
int hcfd_mesh_import_msh(const char* path, hcfd_mesh** out);

#ifdef __cplusplus
}
#endif
#endif /* HCFD_H */
