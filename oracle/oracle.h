/*
 * oracle.h -- TEST INFRASTRUCTURE (the oracle). Not part of the product.
 *
 * A plain-C, OpenMP, FP64 restatement of the reference's PDDSparse CPU path.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it
 * (as oracle/liboracle.so, through oracle/oracle.py); the product library
 * never links it.  Each function cites the reference file:line it restates.
 *
 * Parity status:
 *   orc_philox4x32_10          pinned: published Philox4x32-10 known answers
 *   orc_exp/log/sinpi/...      pinned: libm within a few ulp, mpmath spot checks
 *   orc_rect_dist              pinned: closed-form geometry cases
 *   orc_mesh_*                 pinned: Output/Debug/G.csv pattern (bit-exact),
 *                              Output/solution.csv positions, cond.txt sizes/kinds
 *   orc_stencil_ipsi           pinned: Psi*Psi^-1 = I, cond.txt condition numbers
 *   orc_simulate               pinned: exit-time / harmonic-measure closed forms,
 *                              Gobet bias order, fixture G/B at 3 SE
 *   orc_reduce_faithful        pinned: fixture G/B/G_var at 3 SE; equals
 *                              orc_reduce_batched's G to rounding
 *   orc_bicgstab               pinned: dense LU solve of the same system
 *   orc_lut_*                  pinned: partition of unity, linear reproduction,
 *                              node interpolation, brute-force 16-tap sums
 */
#ifndef PDDS_ORACLE_H
#define PDDS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* PDE identifiers (reference convention 1/2 s s^T : D2u + b.Du + c u + f = 0) */
enum {
    ORC_PDE_MONEGROS = 0,     /* BVPs/EDPs/Monegros_Poisson_CPP.hpp:12-88 */
    ORC_PDE_SINSIN = 1,       /* u = sin(pi x) sin(pi y) + x^2 + y^2, Poisson */
    ORC_PDE_VARCOEF = 2,      /* divergence form, u = e^x cos y, analytic coefficients */
    ORC_PDE_VARCOEF_LUT = 3,  /* same PDE, sigma/b/c from a 5-field B-spline LUT */
    ORC_PDE_HELMHOLTZ = 4,    /* u as SINSIN, c = -k^2 */
    ORC_PDE_TABLE = 5         /* every coefficient from the 8-field LUT (sigma11, sigma22, b1, b2,
                                 c, f, du/dx, du/dy), u = g from the 1-field ulut */
};

/* A B-spline LUT on its own: coefficient array [ny][nx][nf] (float), t = (x - ox) * sx,
   u = (y - oy) * sy */
typedef struct {
    const float *c;
    int nx, ny, nf;
    double ox, oy, sx, sy;
} orc_lut;

typedef struct {
    int pde;
    double k2;              /* HELMHOLTZ: k^2 */
    /* LUT (VARCOEF_LUT): coefficient array [ny][nx][nf] (float), evaluation
       t = (x - ox) * sx, u = (y - oy) * sy; taps clamped to [0,nx-1]x[0,ny-1] */
    const float *lut;
    int lut_nx, lut_ny, lut_nf;
    double lut_ox, lut_oy, lut_sx, lut_sy;
    int step;               /* step family: 0 control variate (Integrators/FKACIntegrator.hpp:46-57),
                               1 plain, grad u taken as 0 (:20-29), 2 VR, the drift change by
                               mu = -sigma^T grad u / u (:31-44), 3 VR+CV with mu and F from
                               vlut (:100-112) */
    /* reflecting sides of the stopping rectangle (bit 0 N, 1 S, 2 E, 3 W): the Lepingle step
       (Integrators/FKACIntegrator.hpp:114-165) with du/dn = robin u + psi, psi from the
       solution; rho: Lepingle step within this distance (0: always) */
    int refl;
    double robin, rho;
    orc_lut ulut;           /* TABLE: u (1 field) */
    orc_lut vlut;           /* step 3: mu1, mu2, F1, F2 */
} orc_problem;

typedef struct {
    double h;               /* Euler-Maruyama step */
    double beta;            /* Gobet constant (reference -0.5826) */
    uint32_t seed;          /* Philox key word 0 */
    uint32_t stream;        /* Philox key word 1 (request id; 0 for assembly) */
    int normals;            /* 0: one Philox call per step, 53-bit uniforms;
                               1: one call per two steps, 24-bit uniforms (FP32 engine streams) */
    int phase2;             /* 1: continue each path to the true exit (beta = 0) */
    int64_t max_steps;      /* per-phase step cap (0: no cap) */
    int threads;            /* OpenMP threads (0: default) */
} orc_run;

/* ---- elementary pieces (for tests) ---- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_math_eval(int fn, const double *x, double *y, int64_t n);   /* 0 exp,1 log,2 sin,3 cos,4 tanh,5 sinpi,6 cospi,7 expm1,8 table exp (x <= 0) */
/* Gaussian RBF values of the m entries of a straight, uniformly spaced stencil side at X, as the
   batched reduction forms them on rectangle sides (chunked two-multiply recurrence) */
void orc_rbf_side_eval(const double *px, const double *py, int m, double x, double y, double ic2, double *phi);
void orc_normals(const orc_run *run, uint32_t knot, uint32_t path, int64_t step0, int64_t nsteps, double *n1, double *n2);
double orc_rect_dist(const double p[4], double x, double y, double *ex, double *ey, double *nx, double *ny);
double orc_sector_dist(const double g[6], double x, double y, double *ex, double *ey, double *nx, double *ny,
                       int *kind);
void orc_pde_eval(const orc_problem *pb, double x, double y, double sig[4], double b[2],
                  double *c, double *f, double grad[2], double *g, double *u);
double orc_lut_eval(const orc_problem *pb, int field, double x, double y);

/* ---- geometry: PDDSparseGM::Fullfill_interfaces / Labels_Stencil / Compute_Stencil ---- */
typedef struct orc_mesh orc_mesh;
orc_mesh *orc_mesh_create(double sw_x, double sw_y, double ne_x, double ne_y, int n_sub, int n_node);
void orc_mesh_destroy(orc_mesh *m);
int orc_mesh_n_knots(const orc_mesh *m);
int orc_mesh_n_interfaces(const orc_mesh *m);
/* knot start positions (dispatch rule), Node_Position positions, boundary flags */
void orc_mesh_knots(const orc_mesh *m, double *x0, double *y0, double *nx, double *ny, int *on_boundary);
/* stencil of knot k: rect[4], sizes m[4], kinds[4] (sides N,S,E,W); if idx != NULL,
   fills idx/px/py with the concatenated side lists (N then S then E then W) */
int orc_mesh_stencil(const orc_mesh *m, int k, double rect[4], int msz[4], int kind[4],
                     int *idx, double *px, double *py);

/* Psi^-1 of one stencil side by full-pivot LU, Gaussian RBF with shape c.
   Returns 0 on success, 1 if singular.  ipsi is row-major m x m. cond = |Psi|_F |Psi^-1|_F */
int orc_stencil_ipsi(int m, const double *px, const double *py, double c, double *ipsi, double *cond);

/* ---- Monte Carlo: EMFKACSolver::Simulate_OMP (phase 1, optional phase 2) ----
   outputs per path (any may be NULL): exit projection, Y, Z, xi, steps;
   phase-2 outputs: projection, Y, Z, xi, steps (only if run->phase2) */
void orc_simulate(const orc_problem *pb, const orc_run *run, double x0, double y0, const double rect[4],
                  uint32_t knot, int64_t path0, int64_t npaths,
                  double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps,
                  double *xt2, double *yt2, double *Y2, double *Z2, double *xi2, int64_t *steps2);
/* phase 1 on a stopping domain of geometry geom: 0 rectangle dom[4], 1 annular sector dom[6] */
void orc_simulate_geom(const orc_problem *pb, const orc_run *run, double x0, double y0, int geom,
                       const double *dom, uint32_t knot, int64_t path0, int64_t npaths,
                       double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps);

/* ---- reductions: Reduce_Analytic + Stencil::G_update + G_return_withrep ----
   Inputs: per-path records, stencil (sizes, concatenated positions/indices, per-side shape c,
   per-side ipsi concatenated row-major), global rect.
   quirk = 1: reproduce the "=+" assignment of Meshes/stencil.cpp:376-395 (last side wins).
   Outputs: CSR row (sorted unique columns) cols[], G[], Gvar[]; returns nnz.
   B, BB (second moment), side counters cnt[6] (N,S,E,W,global,unclassified). */
int orc_reduce_faithful(const orc_problem *pb, int64_t npaths, const double *xt, const double *yt,
                        const double *Y, const double *Z, const double *xi,
                        const double rect[4], const double glob[4], const int msz[4],
                        const int *idx, const double *px, const double *py, const double cshape[4],
                        const double *ipsi, int quirk,
                        int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6]);
/* Restructured estimator: per sub-batch of S paths, Phi = sum Y phi(X) per side, G_beta = -Psi^-T Phi;
   G = sum G_beta / N; Gvar = (sum G_beta^2/n_beta - N G^2) / (nbatches - 1)  (batch means).
   Sums run sub-batch -> item of P paths -> knot, the order the GPU engine uses. */
int orc_reduce_batched(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt, const double *yt,
                       const double *Y, const double *Z, const double *xi,
                       const double rect[4], const double glob[4], const int msz[4],
                       const int *idx, const double *px, const double *py, const double cshape[4],
                       const double *ipsi, int quirk,
                       int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6]);
/* orc_reduce_batched plus the knot equation's mean and batch-means variance with per-column
   reference values w (the engine's pdds_assemble_eq); w may be NULL */
int orc_reduce_batched_eq(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt,
                          const double *yt, const double *Y, const double *Z, const double *xi,
                          const double rect[4], const double glob[4], const int msz[4],
                          const int *idx, const double *px, const double *py, const double cshape[4],
                          const double *ipsi, int quirk, const double *w,
                          int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6],
                          double *eq_mean, double *eq_var);
/* orc_reduce_batched_eq on a stopping domain of geometry geom (disc radius in glob[2]); cent: per-entry
   Gaussian widths (NULL: cshape of the entry's side) */
int orc_reduce_batched_geom(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt,
                            const double *yt, const double *Y, const double *Z, const double *xi, int geom,
                            const double *rect, const double glob[4], const int msz[4],
                            const int *idx, const double *px, const double *py, const double cshape[4],
                            const double *ipsi, int quirk, const double *w, const double *cent,
                            int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6],
                            double *eq_mean, double *eq_var);

/* ---- PDDSparseGM::Compute_Solution: BiCGSTAB on (I + G) u = B, zero guess ---- */
int orc_bicgstab(int n, const int64_t *rowptr, const int *cols, const double *vals, const double *b,
                 double *x, double tol, int maxit, double *relres);

/* ---- LUT construction (host side of BVPs/LUT2.cuh:11-45 and its cubic B-spline semantics) ----
   mode 0: interpolating (prefiltered) coefficients on node-aligned samples, (n+2)^2 output;
   mode 1: raw samples, texel-centre convention (cubicTex2DSimple semantics), n^2 output. */
void orc_lut_prefilter(int nx, int ny, int nf, const double *samples, float *coef);

#ifdef __cplusplus
}
#endif
#endif
