/*
 * hcfd_oracle.c -- CPU oracle for the explicit unstructured finite-volume Euler
 * residual (Green-Gauss gradient, min/max limiter, MUSCL + Rusanov flux,
 * ghost-cell boundary conditions, forward-Euler update).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the CPU restatement that the HIP
 * path is compared against.  Only the test-suite, the smoke entry point and the CPU
 * baseline leg of the benchmark load it; the product library never links or
 * calls it, and it is never the thing that is measured or shipped.
 *
 * It restates, in plain C with the same floating-point expression order, the
 * reference's canonical scalar layout (namespace Base) so that with
 * -ffp-contract=off its results are bit-identical to the reference built the
 * same way.  Each function cites the reference file:line it follows.
 *
 * Array layouts (structure of arrays, like the reference's Base/SoA state):
 *   U, phi, R, Q : [4][Nc]            element (v,i)   at v*Nc + i
 *   G            : [4][2][Nc]         element (v,d,i) at v*2*Nc + d*Nc + i
 *                  (reference grad_u[v][d*Nc+i], includes/Base/GradientSchemes.hpp)
 *   face geometry: rf[Nf][2][2], nf[Nf][2], xf[Nf][2], Af[Nf], If[Nf]
 *   cell geometry: xc[Nc][2], iv[Nc]
 *
 * Extensions that have no reference counterpart (marked NEW below):
 *   - orc_cfl_dt: the CFL time-step formula used by the GPU path;
 *   - orc_run_ssprk: SSP-RK2/RK3 steps built from the forward-Euler stage;
 *   - SubOutlet follows the corrected form of the reference's CAoS layout
 *     (src/CAoS/BoundaryConditions/Outlet.cpp:15-26) because the Base form
 *     reads rul[-1] (src/Base/BoundaryConditions/Outlet.cpp:26-28).
 *
 * Parity pins (tests/test_oracle.py):
 *   geometry, gradient, limiter, BC copies, flux, update, BC cells, prims:
 *     end-to-end golden sols/Base_0.hsln + sols/Base_1.hsln (wedge, 1000 steps)
 *     and bit-exact per-stage comparison with the reference's own sources
 *     compiled under oracle/_ref on wedge and generated meshes; plus closed
 *     forms (free-stream preservation, linear-field gradient, Rusanov
 *     consistency F(U,U)=E(U).n, limiter bounds, conservation).
 *   orc_cfl_dt (NEW): closed form on a uniform right-triangle mesh with a
 *     uniform state, and brute force in Python on tiny meshes.
 *   orc_run_ssprk (NEW): each step equals the textbook combinations of the
 *     reference's own forward-Euler steps (oracle/_ref), bit for bit; its
 *     temporal order (errors falling as h^2 and h^3) on a smooth flow.
 *   SubInlet/SubOutlet/Wall: closed-form ghost states in the tests (the
 *     golden case only exercises Symmetry, SupInlet and SupOutlet).
 *   orc_initialise: sols/Base_0.hsln.  orc_initialise_prims (the per-cell
 *     initial state): a uniform state gives orc_initialise's bits, a piecewise
 *     state gives each interior cell the bits of the uniform form of its value
 *     and the ghosts the BC update, permuting the state over the interior
 *     permutes the result exactly, and orc_prims returns it to a few ulps.
 *   The whole file also runs under AddressSanitizer and
 *     UndefinedBehaviorSanitizer (tests/test_sanitize.py) with its ordinary
 *     build's results bit for bit.
 *   orc_rusanov3d / orc_euler_flux3d / orc_prims3d / orc_cons3d (the NDIMS == 3
 *     branches, which no mesh of the reference's format can run): with w = 0
 *     and n_z = 0 they equal the 2-D functions bit for bit (pinned above);
 *     closed-form Euler flux, Rusanov consistency and antisymmetry, rotation
 *     invariance, the prim <-> cons round trip.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 4
#define ND 2

enum { BC_SYMMETRY = 0, BC_WALL, BC_SUPINLET, BC_SUBINLET, BC_SUPOUTLET, BC_SUBOUTLET };

typedef struct orc_dom {
    int Nc, Nf, Ncol, NBC, F_BC;
    int* coloff;            /* [Ncol+1] */
    int* C;                 /* [Nf][2] */
    double* rf;             /* [Nf][2][2] */
    double* nf;             /* [Nf][2] */
    double* xf;             /* [Nf][2] */
    double* Af;             /* [Nf] */
    double* If;             /* [Nf] */
    double* xc;             /* [Nc][2] */
    double* iv;             /* [Nc] */
    int* cf_off;            /* [Nc+1] cell->face CSR, ascending face id */
    int* cf_con;            /* [2Nf] */
    int bc_iv_zeroed;
} orc_dom;

static double dot2(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1]; }

/* Mesh::compute_cf_con, src/Mesh.cpp:88-117 (count, scan, fill in face order). */
static void build_cf(orc_dom* d) {
    int Nc = d->Nc, Nf = d->Nf;
    d->cf_off = (int*)calloc((size_t)Nc + 1, sizeof(int));
    for (int f = 0; f < Nf; f++) {
        d->cf_off[d->C[2 * f] + 1]++;
        d->cf_off[d->C[2 * f + 1] + 1]++;
    }
    for (int i = 0; i < Nc; i++) d->cf_off[i + 1] += d->cf_off[i];
    d->cf_con = (int*)calloc((size_t)d->cf_off[Nc] + 1, sizeof(int));
    int* cnt = (int*)calloc((size_t)Nc, sizeof(int));
    for (int f = 0; f < Nf; f++) {
        int CL = d->C[2 * f], CR = d->C[2 * f + 1];
        d->cf_con[d->cf_off[CL] + cnt[CL]++] = f;
        d->cf_con[d->cf_off[CR] + cnt[CR]++] = f;
    }
    free(cnt);
}

/* Base::Domain::Domain and its compute_* helpers, src/Base/Domain.cpp:4-146. */
orc_dom* orc_dom_create(int Nn, int Nc, int Nf, int Ncol, int NBCcol, const int* coloff,
                        const double* nodes, const int* fn_con, const int* fc_con,
                        const int* cn_off, const int* cn_con) {
    (void)Nn;
    orc_dom* d = (orc_dom*)calloc(1, sizeof(orc_dom));
    d->Nc = Nc; d->Nf = Nf; d->Ncol = Ncol; d->NBC = NBCcol;
    d->coloff = (int*)malloc(sizeof(int) * (size_t)(Ncol + 1));
    memcpy(d->coloff, coloff, sizeof(int) * (size_t)(Ncol + 1));
    d->F_BC = coloff[Ncol - NBCcol];
    d->C  = (int*)calloc((size_t)Nf * 2, sizeof(int));
    d->rf = (double*)calloc((size_t)Nf * 4, sizeof(double));
    d->nf = (double*)calloc((size_t)Nf * 2, sizeof(double));
    d->xf = (double*)calloc((size_t)Nf * 2, sizeof(double));
    d->Af = (double*)calloc((size_t)Nf, sizeof(double));
    d->If = (double*)calloc((size_t)Nf, sizeof(double));
    d->xc = (double*)calloc((size_t)Nc * 2, sizeof(double));
    d->iv = (double*)calloc((size_t)Nc, sizeof(double));

    /* get_face_con :37-42 */
    for (int f = 0; f < Nf; f++) { d->C[2 * f] = fc_con[2 * f]; d->C[2 * f + 1] = fc_con[2 * f + 1]; }
    /* compute_xf :44-59 */
    for (int f = 0; f < Nf; f++) {
        int n1 = fn_con[2 * f], n2 = fn_con[2 * f + 1];
        for (int k = 0; k < ND; k++) d->xf[2 * f + k] = 0.5 * (nodes[ND * n1 + k] + nodes[ND * n2 + k]);
    }
    /* compute_xc :61-71 -- node mean accumulated in cn_con order */
    for (int i = 0; i < Nc; i++) {
        int nnds = cn_off[i + 1] - cn_off[i];
        for (int n = cn_off[i]; n < cn_off[i + 1]; n++) {
            int nd = cn_con[n];
            for (int k = 0; k < ND; k++) d->xc[2 * i + k] += nodes[nd * ND + k] / nnds;
        }
    }
    /* compute_rf :73-89 -- BC faces get rR = -rL */
    for (int f = 0; f < Nf; f++) {
        int CL = d->C[2 * f], CR = d->C[2 * f + 1];
        for (int k = 0; k < ND; k++) {
            d->rf[4 * f + 0 + k] = d->xf[2 * f + k] - d->xc[2 * CL + k];
            d->rf[4 * f + 2 + k] = d->xf[2 * f + k] - d->xc[2 * CR + k];
        }
    }
    for (int f = d->F_BC; f < Nf; f++)
        for (int k = 0; k < ND; k++) d->rf[4 * f + 2 + k] = -d->rf[4 * f + 0 + k];
    /* compute_nf :91-116 -- unit normal oriented from L to R */
    for (int f = 0; f < Nf; f++) {
        int n1 = fn_con[2 * f], n2 = fn_con[2 * f + 1];
        double* n = &d->nf[2 * f];
        n[0] = (nodes[ND * n2 + 1] - nodes[ND * n1 + 1]);
        n[1] = -(nodes[ND * n2 + 0] - nodes[ND * n1 + 0]);
        d->Af[f] = sqrt(dot2(n, n));
        for (int k = 0; k < ND; k++) n[k] /= d->Af[f];
        double ndr = dot2(n, &d->rf[4 * f]);
        int sgn = (ndr > 0) ? 1 : -1;
        n[0] = sgn * n[0];
        n[1] = sgn * n[1];
    }
    /* compute_If :118-125 */
    for (int f = 0; f < Nf; f++) {
        double nrL = dot2(&d->nf[2 * f], &d->rf[4 * f + 0]);
        double nrR = -dot2(&d->nf[2 * f], &d->rf[4 * f + 2]);
        d->If[f] = nrL / (nrL + nrR);
    }
    /* compute_iv :127-146 */
    for (int f = 0; f < Nf; f++) {
        double aL = d->Af[f] * dot2(&d->nf[2 * f], &d->rf[4 * f + 0]);
        double aR = d->Af[f] * dot2(&d->nf[2 * f], &d->rf[4 * f + 2]);
        int CL = d->C[2 * f], CR = d->C[2 * f + 1];
        d->iv[CL] += aL / ND;
        d->iv[CR] -= aR / ND;
    }
    for (int i = 0; i < Nc; i++) d->iv[i] = 1.0 / d->iv[i];
    for (int f = d->F_BC; f < Nf; f++) d->iv[d->C[2 * f + 1]] = d->iv[d->C[2 * f]];

    build_cf(d);
    return d;
}

void orc_dom_free(orc_dom* d) {
    if (!d) return;
    free(d->coloff); free(d->C); free(d->rf); free(d->nf); free(d->xf); free(d->Af); free(d->If);
    free(d->xc); free(d->iv); free(d->cf_off); free(d->cf_con); free(d);
}

/* Base::Domain::zero_bc_iv, src/Base/Domain.cpp:148-154 (called by the Euler ctor, Euler.cpp:24). */
void orc_zero_bc_iv(orc_dom* d) {
    for (int f = d->F_BC; f < d->Nf; f++) d->iv[d->C[2 * f + 1]] = 0.0;
    d->bc_iv_zeroed = 1;
}

/* Copies of the geometry (any pointer may be NULL). */
void orc_geometry(const orc_dom* d, int* C, double* xf, double* xc, double* rf, double* nf,
                  double* Af, double* If, double* iv) {
    size_t Nf = (size_t)d->Nf, Nc = (size_t)d->Nc;
    if (C) memcpy(C, d->C, sizeof(int) * Nf * 2);
    if (xf) memcpy(xf, d->xf, sizeof(double) * Nf * 2);
    if (xc) memcpy(xc, d->xc, sizeof(double) * Nc * 2);
    if (rf) memcpy(rf, d->rf, sizeof(double) * Nf * 4);
    if (nf) memcpy(nf, d->nf, sizeof(double) * Nf * 2);
    if (Af) memcpy(Af, d->Af, sizeof(double) * Nf);
    if (If) memcpy(If, d->If, sizeof(double) * Nf);
    if (iv) memcpy(iv, d->iv, sizeof(double) * Nc);
}

void orc_cf(const orc_dom* d, int* cf_off, int* cf_con) {
    memcpy(cf_off, d->cf_off, sizeof(int) * ((size_t)d->Nc + 1));
    memcpy(cf_con, d->cf_con, sizeof(int) * (size_t)d->cf_off[d->Nc]);
}

/* GaussGreen::compute_gradient, src/Base/GaussGreen.cpp:14-40. */
void orc_gradient(const orc_dom* d, const double* U, double* G) {
    int Nc = d->Nc, Nf = d->Nf;
    memset(G, 0, sizeof(double) * (size_t)NV * ND * Nc);
    for (int f = 0; f < Nf; f++) {
        int CL = d->C[2 * f], CR = d->C[2 * f + 1];
        double If = d->If[f], Af = d->Af[f];
        for (int v = 0; v < NV; v++) {
            const double* u = U + (size_t)v * Nc;
            double* g = G + (size_t)v * ND * Nc;
            for (int k = 0; k < ND; k++) {
                double uf = u[CL] * If + u[CR] * (1 - If);
                g[k * Nc + CL] += uf * Af * d->nf[2 * f + k];
                g[k * Nc + CR] -= uf * Af * d->nf[2 * f + k];
            }
        }
    }
    for (int v = 0; v < NV; v++)
        for (int k = 0; k < ND; k++)
            for (int i = 0; i < Nc; i++) G[(size_t)v * ND * Nc + (size_t)k * Nc + i] *= d->iv[i];
}

/* Limiter::limit_gradient, src/Base/Limiter.cpp:19-77. */
void orc_limiter(const orc_dom* d, const double* U, const double* G, double* phi) {
    const double eps = 1e-9;
    int Nc = d->Nc, Nf = d->Nf;
    double* Umin = (double*)malloc(sizeof(double) * (size_t)Nc);
    double* Umax = (double*)malloc(sizeof(double) * (size_t)Nc);
    double* dL = (double*)malloc(sizeof(double) * (size_t)Nf);
    double* dR = (double*)malloc(sizeof(double) * (size_t)Nf);
    for (int v = 0; v < NV; v++) {
        const double* u = U + (size_t)v * Nc;
        const double* g = G + (size_t)v * ND * Nc;
        double* p = phi + (size_t)v * Nc;
        memcpy(Umin, u, sizeof(double) * (size_t)Nc);
        memcpy(Umax, u, sizeof(double) * (size_t)Nc);
        for (int i = 0; i < Nc; i++) p[i] = 1.0;
        for (int f = 0; f < Nf; f++) {
            int CL = d->C[2 * f], CR = d->C[2 * f + 1];
            Umin[CL] = fmin(Umin[CL], u[CR]);
            Umax[CL] = fmax(Umax[CL], u[CR]);
            Umin[CR] = fmin(Umin[CR], u[CL]);
            Umax[CR] = fmax(Umax[CR], u[CL]);
        }
        for (int f = 0; f < Nf; f++) {
            int CL = d->C[2 * f], CR = d->C[2 * f + 1];
            double GL[ND] = {g[CL], g[Nc + CL]};
            double GR[ND] = {g[CR], g[Nc + CR]};
            dL[f] = dot2(GL, &d->rf[4 * f + 0]);
            dR[f] = dot2(GR, &d->rf[4 * f + 2]);
        }
        for (int f = 0; f < Nf; f++) {
            int CL = d->C[2 * f], CR = d->C[2 * f + 1];
            double UL = u[CL], UR = u[CR];
            double pl = 1.0, pr = 1.0;
            double a = dL[f], b = dR[f];
            pl = (a > eps) ? fmin(1, (Umax[CL] - UL) / a) : pl;
            pl = (a < -eps) ? fmin(1, (Umin[CL] - UL) / a) : pl;
            pr = (b > eps) ? fmin(1, (Umax[CR] - UR) / b) : pr;
            pr = (b < -eps) ? fmin(1, (Umin[CR] - UR) / b) : pr;
            p[CL] = fmin(p[CL], pl);
            p[CR] = fmin(p[CR], pr);
        }
    }
    free(Umin); free(Umax); free(dL); free(dR);
}

/* BoundaryCondition::update_grads / update_limiter, src/Base/BoundaryConditions/Base.cpp:29-53,
 * applied for every BC in colour order (Euler::update_BC_grads/limiter, Euler.cpp:144-154). */
void orc_bc_grads(const orc_dom* d, double* G) {
    int Nc = d->Nc;
    for (int v = 0; v < NV; v++)
        for (int k = 0; k < ND; k++)
            for (int f = d->F_BC; f < d->Nf; f++) {
                int CL = d->C[2 * f], CR = d->C[2 * f + 1];
                G[(size_t)v * ND * Nc + (size_t)k * Nc + CR] = -G[(size_t)v * ND * Nc + (size_t)k * Nc + CL];
            }
}

void orc_bc_limiter(const orc_dom* d, double* phi) {
    int Nc = d->Nc;
    for (int v = 0; v < NV; v++)
        for (int f = d->F_BC; f < d->Nf; f++) {
            int CL = d->C[2 * f], CR = d->C[2 * f + 1];
            phi[(size_t)v * Nc + CR] = phi[(size_t)v * Nc + CL];
        }
}

/* Flux::EulerFlux, src/Base/Flux.cpp:89-97. */
static void euler_flux(double gma, const double U[NV], const double n[ND], double F[NV]) {
    double gmo = gma - 1;
    double ir = 1.0 / U[0];
    double P = gmo * (U[3] - 0.5 * ir * dot2(&U[1], &U[1]));
    F[0] = n[0] * U[1] + n[1] * U[2];
    F[1] = n[0] * (U[1] * U[1] * ir + P) + n[1] * U[2] * U[1] * ir;
    F[2] = n[0] * U[1] * U[2] * ir + n[1] * (U[2] * U[2] * ir + P);
    F[3] = n[0] * (U[3] + P) * U[1] * ir + n[1] * (U[3] + P) * U[2] * ir;
}

/* Flux::Rusanov, src/Base/Flux.cpp:52-81. */
static void rusanov(double gma, const double UL[NV], const double UR[NV], const double n[ND], double F[NV]) {
    double gmo = gma - 1;
    double irl = 1.0 / UL[0];
    double irr = 1.0 / UR[0];
    double nvl = dot2(n, &UL[1]) * irl;
    double nvr = dot2(n, &UR[1]) * irr;
    double qql = 0.5 * dot2(&UL[1], &UL[1]) * irl;
    double qqr = 0.5 * dot2(&UR[1], &UR[1]) * irr;
    double pl = gmo * (UL[3] - qql);
    double pr = gmo * (UR[3] - qqr);
    double al = sqrt(gma * pl * irl);
    double ar = sqrt(gma * pr * irr);
    double S1 = fmax(fabs(nvl - al), fabs(nvr - ar));
    double S2 = fmax(fabs(nvl + al), fabs(nvr + ar));
    double S = fmax(S1, S2);
    double FL[NV], FR[NV];
    euler_flux(gma, UL, n, FL);
    euler_flux(gma, UR, n, FR);
    for (int v = 0; v < NV; v++) F[v] = 0.5 * ((FL[v] + FR[v]) - S * (UR[v] - UL[v]));
}

/* Single-face entry for closed-form tests: F = Rusanov(UL, UR, n). */
void orc_rusanov(double gma, const double* UL, const double* UR, const double* n, double* F) {
    rusanov(gma, UL, UR, n, F);
}

void orc_euler_flux(double gma, const double* U, const double* n, double* F) { euler_flux(gma, U, n, F); }

/* Flux::ComputeFlux with MUSCL, src/Base/Flux.cpp:9-50 and :83-87.  R accumulates. */
void orc_flux(const orc_dom* d, double gma, const double* U, const double* G, const double* phi, double* R) {
    int Nc = d->Nc, Nf = d->Nf;
    double UL[NV], UR[NV], UfL[NV], UfR[NV], F[NV];
    for (int f = 0; f < Nf; f++) {
        int CL = d->C[2 * f], CR = d->C[2 * f + 1];
        double Af = d->Af[f];
        const double* rL = &d->rf[4 * f + 0];
        const double* rR = &d->rf[4 * f + 2];
        for (int v = 0; v < NV; v++) {
            UL[v] = U[(size_t)v * Nc + CL];
            UR[v] = U[(size_t)v * Nc + CR];
            double GL[ND] = {G[(size_t)v * ND * Nc + CL], G[(size_t)v * ND * Nc + Nc + CL]};
            double GR[ND] = {G[(size_t)v * ND * Nc + CR], G[(size_t)v * ND * Nc + Nc + CR]};
            UfL[v] = UL[v] + phi[(size_t)v * Nc + CL] * dot2(GL, rL);
            UfR[v] = UR[v] + phi[(size_t)v * Nc + CR] * dot2(GR, rR);
        }
        rusanov(gma, UfL, UfR, &d->nf[2 * f], F);
        for (int v = 0; v < NV; v++) {
            R[(size_t)v * Nc + CL] -= Af * F[v];
            R[(size_t)v * Nc + CR] += Af * F[v];
        }
    }
    for (int v = 0; v < NV; v++)
        for (int i = 0; i < Nc; i++) R[(size_t)v * Nc + i] *= d->iv[i];
}

/* Euler::add_residuals + residual clear, src/Base/Euler.cpp:107-113 and :93-95. */
void orc_update(const orc_dom* d, double dt, double* U, double* R) {
    size_t n = (size_t)NV * d->Nc;
    for (size_t k = 0; k < n; k++) U[k] += dt * R[k];
    memset(R, 0, sizeof(double) * n);
}

/* BoundaryCondition::init's UBC, src/Base/BoundaryConditions/Base.cpp:18-22. */
static void cons_of_prims(double gma, const double Q[NV], double Uo[NV]) {
    Uo[0] = Q[0];
    for (int v = 1; v < NV; v++) Uo[v] = Q[0] * Q[v];
    Uo[3] = Q[3] / (gma - 1) + 0.5 * Q[0] * dot2(&Q[1], &Q[1]);
}

void orc_cons_of_prims(double gma, const double* Q, double* Uo) { cons_of_prims(gma, Q, Uo); }

/* BC update_cells in BC order: Symmetry.cpp:5-24, Wall.cpp:4-15, Inlet.cpp:4-31,
 * Outlet.cpp:4-13; SubOutlet per the corrected CAoS form (CAoS Outlet.cpp:15-26). */
void orc_bc_cells(const orc_dom* d, const int* bctype, const double* qbc, double gma, double* U) {
    int Nc = d->Nc;
    double* u0 = U; double* u1 = U + Nc; double* u2 = U + 2 * (size_t)Nc; double* u3 = U + 3 * (size_t)Nc;
    for (int bc = 0; bc < d->NBC; bc++) {
        int fcol = d->Ncol - d->NBC + bc;
        int F1 = d->coloff[fcol], F2 = d->coloff[fcol + 1];
        const double* Q = qbc + NV * bc;
        double UBC[NV];
        cons_of_prims(gma, Q, UBC);
        for (int f = F1; f < F2; f++) {
            int CL = d->C[2 * f], CR = d->C[2 * f + 1];
            switch (bctype[bc]) {
            case BC_SYMMETRY: {
                const double* n = &d->nf[2 * f];
                double ul[ND] = {u1[CL], u2[CL]};
                double nor = dot2(ul, n);
                u0[CR] = u0[CL];
                u1[CR] = ul[0] - 2 * nor * n[0];
                u2[CR] = ul[1] - 2 * nor * n[1];
                u3[CR] = u3[CL];
            } break;
            case BC_WALL:
                u0[CR] = u0[CL]; u1[CR] = -u1[CL]; u2[CR] = -u2[CL]; u3[CR] = u3[CL];
                break;
            case BC_SUPINLET:
                u0[CR] = UBC[0]; u1[CR] = UBC[1]; u2[CR] = UBC[2]; u3[CR] = UBC[3];
                break;
            case BC_SUBINLET: {
                double UL[NV] = {u0[CL], u1[CL], u2[CL], u3[CL]};
                double PL = (gma - 1) * (UL[3] - 0.5 * dot2(&UL[1], &UL[1]) / UL[0]);
                u0[CR] = UBC[0]; u1[CR] = UBC[1]; u2[CR] = UBC[2]; u3[CR] = UBC[3];
                u3[CR] = PL / (gma - 1) + 0.5 * Q[0] * dot2(&Q[1], &Q[1]);
            } break;
            case BC_SUPOUTLET:
                u0[CR] = u0[CL]; u1[CR] = u1[CL]; u2[CR] = u2[CL]; u3[CR] = u3[CL];
                break;
            case BC_SUBOUTLET: {
                double ul[NV] = {u0[CL], u1[CL], u2[CL], u3[CL]};
                u0[CR] = ul[0]; u1[CR] = ul[1]; u2[CR] = ul[2]; u3[CR] = ul[3];
                u3[CR] = Q[3] / (gma - 1) + 0.5 * dot2(&ul[1], &ul[1]) / ul[0];
            } break;
            default: break;
            }
        }
    }
}

/* Euler::initialise, src/Base/Euler.cpp:62-77: uniform state then BC ghosts. */
void orc_initialise(const orc_dom* d, const double* QIC, const int* bctype, const double* qbc, double gma,
                    double* U) {
    double UIC[NV];
    UIC[0] = QIC[0];
    for (int v = 1; v < NV; v++) UIC[v] = QIC[0] * QIC[v];
    UIC[3] = QIC[3] / (gma - 1) + 0.5 * QIC[0] * dot2(&QIC[1], &QIC[1]);
    for (int v = 0; v < NV; v++)
        for (int i = 0; i < d->Nc; i++) U[(size_t)v * d->Nc + i] = UIC[v];
    orc_bc_cells(d, bctype, qbc, gma, U);
}

/* Per-cell primitive initial condition (same arithmetic as Euler::initialise,
 * applied cell by cell), followed by the BC ghost update. */
void orc_initialise_prims(const orc_dom* d, const double* Q, const int* bctype, const double* qbc, double gma,
                          double* U) {
    int Nc = d->Nc;
    for (int i = 0; i < Nc; i++) {
        double q[NV] = {Q[i], Q[Nc + i], Q[2 * (size_t)Nc + i], Q[3 * (size_t)Nc + i]};
        double u[NV];
        u[0] = q[0];
        for (int v = 1; v < NV; v++) u[v] = q[0] * q[v];
        u[3] = q[3] / (gma - 1) + 0.5 * q[0] * dot2(&q[1], &q[1]);
        for (int v = 0; v < NV; v++) U[(size_t)v * Nc + i] = u[v];
    }
    orc_bc_cells(d, bctype, qbc, gma, U);
}

/* Euler::compute_prims, src/Base/Euler.cpp:115-136. */
void orc_prims(const orc_dom* d, double gma, const double* U, double* Q) {
    int Nc = d->Nc;
    for (int i = 0; i < Nc; i++) Q[i] = U[i];
    for (int v = 1; v < NV; v++) {
        if (v == ND + 1) continue;
        for (int i = 0; i < Nc; i++) Q[(size_t)v * Nc + i] = U[(size_t)v * Nc + i] / U[i];
    }
    double gmo = gma - 1;
    for (int i = 0; i < Nc; i++) {
        double u[ND] = {Q[Nc + i], Q[2 * (size_t)Nc + i]};
        Q[3 * (size_t)Nc + i] = gmo * (U[3 * (size_t)Nc + i] - 0.5 * Q[i] * dot2(u, u));
    }
}

/* NEW (no reference counterpart): CFL time step.
 *   lam_i = sum over the cell's faces, ascending face id, of A_f*(|u_i.n_f| + c_i)
 *   dt    = cfl * min over interior cells (iv_i != 0) of 1/(iv_i*lam_i)
 * with u = (U1/U0, U2/U0), p = (g-1)(U3 - 0.5(U1 ux + U2 uy)), c = sqrt(g p / U0). */
double orc_cfl_dt(const orc_dom* d, double gma, double cfl, const double* U) {
    int Nc = d->Nc;
    double gmo = gma - 1;
    double best = INFINITY;
    for (int i = 0; i < Nc; i++) {
        if (d->iv[i] == 0.0) continue;
        double ir = 1.0 / U[i];
        double ux = U[Nc + i] * ir;
        double uy = U[2 * (size_t)Nc + i] * ir;
        double p = gmo * (U[3 * (size_t)Nc + i] - 0.5 * (U[Nc + i] * ux + U[2 * (size_t)Nc + i] * uy));
        double c = sqrt(gma * p * ir);
        double lam = 0.0;
        for (int k = d->cf_off[i]; k < d->cf_off[i + 1]; k++) {
            int f = d->cf_con[k];
            lam += d->Af[f] * (fabs(ux * d->nf[2 * f] + uy * d->nf[2 * f + 1]) + c);
        }
        double dti = 1.0 / (d->iv[i] * lam);
        if (dti < best) best = dti;
    }
    return cfl * best;
}

/* Euler::run's step loop, src/Base/Euler.cpp:79-105 (output omitted): per step
 * gradient, limiter, BC gradient/limiter copies, flux, update, BC cells.
 * cfl > 0 selects the CFL time step (NEW) recomputed at the start of each step;
 * otherwise the fixed dt is used.  Returns the simulated time. */
double orc_run(orc_dom* d, const int* bctype, const double* qbc, double gma, double* U, int nsteps, double dt,
               double cfl) {
    int Nc = d->Nc;
    if (!d->bc_iv_zeroed) orc_zero_bc_iv(d);
    double* G = (double*)calloc((size_t)NV * ND * Nc, sizeof(double));
    double* phi = (double*)calloc((size_t)NV * Nc, sizeof(double));
    double* R = (double*)calloc((size_t)NV * Nc, sizeof(double));
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double h = (cfl > 0) ? orc_cfl_dt(d, gma, cfl, U) : dt;
        orc_gradient(d, U, G);
        orc_limiter(d, U, G, phi);
        orc_bc_grads(d, G);
        orc_bc_limiter(d, phi);
        orc_flux(d, gma, U, G, phi, R);
        orc_update(d, h, U, R);
        orc_bc_cells(d, bctype, qbc, gma, U);
        t += h;
    }
    free(G); free(phi); free(R);
    return t;
}

/* NEW: strong-stability-preserving Runge-Kutta steps (Shu and Osher, J. Comput.
 * Phys. 77 (1988) 439-471; Gottlieb, Shu and Tadmor, SIAM Rev. 43 (2001) 89-112)
 * built from the forward-Euler stage above: with E(W) = W + h R(W) (the update of
 * orc_run, ghost cells then set from the interior),
 *   stages 1: U^{n+1} = E(U^n)                                  (forward Euler)
 *   stages 2: U1 = E(U^n);  U^{n+1} = 1/2 U^n + 1/2 E(U1)             (SSP-RK2)
 *   stages 3: U1 = E(U^n);  U2 = 3/4 U^n + 1/4 E(U1);
 *             U^{n+1} = 1/3 U^n + 2/3 E(U2)                           (SSP-RK3)
 * each combination evaluated as a*U^n + b*E with the two products rounded
 * separately.  h is the fixed dt, or the CFL step of U^n when cfl > 0; the
 * reference has forward Euler only (includes/Base/TimeIntegrators.hpp:11-45).
 * Returns the simulated time, or -1 for an unsupported stage count. */
double orc_run_ssprk(orc_dom* d, const int* bctype, const double* qbc, double gma, double* U, int nsteps,
                     double dt, double cfl, int stages) {
    static const double ca[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
    static const double cb[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
    if (stages < 1 || stages > 3) return -1.0;
    int Nc = d->Nc;
    size_t n = (size_t)NV * Nc;
    if (!d->bc_iv_zeroed) orc_zero_bc_iv(d);
    double* G = (double*)calloc((size_t)NV * ND * Nc, sizeof(double));
    double* phi = (double*)calloc(n, sizeof(double));
    double* R = (double*)calloc(n, sizeof(double));
    double* U0 = (double*)malloc(sizeof(double) * n);
    double t = 0.0;
    for (int s = 0; s < nsteps; s++) {
        double h = (cfl > 0) ? orc_cfl_dt(d, gma, cfl, U) : dt;
        memcpy(U0, U, sizeof(double) * n);
        for (int k = 0; k < stages; k++) {
            orc_gradient(d, U, G);
            orc_limiter(d, U, G, phi);
            orc_bc_grads(d, G);
            orc_bc_limiter(d, phi);
            orc_flux(d, gma, U, G, phi, R);
            for (size_t q = 0; q < n; q++) {
                double e = U[q] + h * R[q];
                U[q] = (k == 0) ? e : ca[stages][k] * U0[q] + cb[stages][k] * e;
            }
            memset(R, 0, sizeof(double) * n);
            orc_bc_cells(d, bctype, qbc, gma, U);
        }
        t += h;
    }
    free(G); free(phi); free(R); free(U0);
    return t;
}

/* ---------------------------------------------------------------------------
 * The NDIMS == 3 branches (includes/Global.hpp:20-21 switch the reference to
 * five variables): state (rho, rho u, rho v, rho w, E), normals of three
 * components.  The reference's Domain builds normals from two-node faces, so
 * these are the flux and state conversions alone, on given arrays.
 * ------------------------------------------------------------------------- */
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Flux::EulerFlux, src/Base/Flux.cpp:89-92, 98-104 (NDIMS == 3). */
static void euler_flux3(double gma, const double U[5], const double n[3], double F[5]) {
    double gmo = gma - 1;
    double ir = 1.0 / U[0];
    double P = gmo * (U[4] - 0.5 * ir * dot3(&U[1], &U[1]));
    F[0] = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
    F[1] = n[0] * (U[1] * U[1] * ir + P) + n[1] * U[2] * U[1] * ir + n[2] * U[3] * U[1] * ir;
    F[2] = n[0] * U[1] * U[2] * ir + n[1] * (U[2] * U[2] * ir + P) + n[2] * U[3] * U[2] * ir;
    F[3] = n[0] * U[1] * U[3] * ir + n[1] * U[2] * U[3] * ir + n[2] * (U[3] * U[3] * ir + P);
    F[4] = n[0] * (U[4] + P) * U[1] * ir + n[1] * (U[4] + P) * U[2] * ir + n[2] * (U[4] + P) * U[3] * ir;
}

/* Flux::Rusanov, src/Base/Flux.cpp:52-81, with NDIMS == 3 dot products. */
static void rusanov3(double gma, const double UL[5], const double UR[5], const double n[3], double F[5]) {
    double gmo = gma - 1;
    double irl = 1.0 / UL[0];
    double irr = 1.0 / UR[0];
    double nvl = dot3(n, &UL[1]) * irl;
    double nvr = dot3(n, &UR[1]) * irr;
    double qql = 0.5 * dot3(&UL[1], &UL[1]) * irl;
    double qqr = 0.5 * dot3(&UR[1], &UR[1]) * irr;
    double pl = gmo * (UL[4] - qql);
    double pr = gmo * (UR[4] - qqr);
    double al = sqrt(gma * pl * irl);
    double ar = sqrt(gma * pr * irr);
    double S1 = fmax(fabs(nvl - al), fabs(nvr - ar));
    double S2 = fmax(fabs(nvl + al), fabs(nvr + ar));
    double S = fmax(S1, S2);
    double FL[5], FR[5];
    euler_flux3(gma, UL, n, FL);
    euler_flux3(gma, UR, n, FR);
    for (int v = 0; v < 5; v++) F[v] = 0.5 * ((FL[v] + FR[v]) - S * (UR[v] - UL[v]));
}

/* n face-state pairs, AoS [n][5] states and fluxes, [n][3] normals. */
void orc_rusanov3d(int n, double gma, const double* UL, const double* UR, const double* nrm, double* F) {
    for (int f = 0; f < n; f++) rusanov3(gma, UL + 5 * (size_t)f, UR + 5 * (size_t)f, nrm + 3 * (size_t)f, F + 5 * (size_t)f);
}

void orc_euler_flux3d(int n, double gma, const double* U, const double* nrm, double* F) {
    for (int f = 0; f < n; f++) euler_flux3(gma, U + 5 * (size_t)f, nrm + 3 * (size_t)f, F + 5 * (size_t)f);
}

/* Euler::compute_prims, src/Base/Euler.cpp:115-136, NDIMS == 3, per state [n][5]. */
void orc_prims3d(int n, double gma, const double* U, double* Q) {
    double gmo = gma - 1;
    for (int i = 0; i < n; i++) {
        const double* u = U + 5 * (size_t)i;
        double* q = Q + 5 * (size_t)i;
        q[0] = u[0];
        for (int v = 1; v < 4; v++) q[v] = u[v] / u[0];
        double w[3] = {q[1], q[2], q[3]};
        q[4] = gmo * (u[4] - 0.5 * q[0] * dot3(w, w));
    }
}

/* Euler::initialise's state conversion, src/Base/Euler.cpp:62-68, NDIMS == 3. */
void orc_cons3d(int n, double gma, const double* Q, double* U) {
    for (int i = 0; i < n; i++) {
        const double* q = Q + 5 * (size_t)i;
        double* u = U + 5 * (size_t)i;
        u[0] = q[0];
        for (int v = 1; v < 5; v++) u[v] = q[0] * q[v];
        u[4] = q[4] / (gma - 1) + 0.5 * q[0] * dot3(&q[1], &q[1]);
    }
}

/* Number of steps Euler::run(T, dt, dt_out) takes: t += dt while t < T
 * (src/Base/Euler.cpp:83-97). */
int orc_count_steps(double T, double dt) {
    double t = 0.0;
    int n = 0;
    while (t < T) { t += dt; n++; }
    return n;
}
