/*
 * pdds_engine.h -- C ABI of the MI355X PDDSparse path engine (libpdds.so).
 *
 * Plain C types only.  Every entry point returns a pdds_status; no entry point
 * prints or exits.  The caller owns every host buffer; the engine owns device
 * memory.  One engine is bound to one HIP device.  Calls on one engine are
 * serialised by the engine; separate engines may be used from separate threads.
 * Results are a function of (seed, stream, knot id, path index) only, never of
 * the launch shape or of what else runs on the device.
 *
 * What each group of entry points replaces in the reference (file:line):
 *   pdds_problem            struct bvp function-pointer plugin (BVPs/BVP.hpp:89-137),
 *                           compile-time device twin (BVPs/BVP.cuh:44-77, BVPs/initBVPdev.cuh:6-16)
 *   pdds_lut_*              Init_tex_LUT / Init_LUT (BVPs/LUT2.cuh:11-92) + cubicTex2DSimple
 *                           (Integrators/cubicTex2D.cu:64-87)
 *   pdds_mesh_*             PDDSparseGM::Fullfill_interfaces (PDDAlgorithms/PDDSparseGM.cpp:346-689),
 *                           Labels_Stencil / Compute_Stencil (:1330-1433, :1845-2041),
 *                           Stencil::Init / Compute_ipsi (Meshes/stencil.cpp:11-85, :115-243)
 *   pdds_paths              MCinCUDA (Solvers/MCinCUDA.cu:441-480, declared Solvers/EMFKACSolver.hpp:10-12):
 *                           per-path exit records, phase 1 and optional phase 2
 *   pdds_solve_knot         EMFKACSolver::Solve_OMP_Analytic with a stencil (Solvers/EMFKACSolver.hpp:711-729):
 *                           one knot's G row, G variances, B and B second moment
 *   pdds_point_eval         EMFKACSolver::Solve_OMP_Analytic without a stencil (Solvers/EMFKACSolver.hpp:686-691):
 *                           Feynman-Kac estimate of u at one point
 *   pdds_assemble           the worker loop of PDDSparseGM::Solve (PDDAlgorithms/PDDSparseGM.cpp:795-848):
 *                           every knot of a mesh in one persistent launch, straight into CSR
 *   pdds_solve              Compute_Solution's BiCGSTAB on (I+G)u = B (PDDAlgorithms/PDDSparseGM.cpp:3168-3246)
 *   pdds_server_*           no reference counterpart: continuous-batching request loop
 */
#ifndef PDDS_ENGINE_H
#define PDDS_ENGINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pdds_status;
#define PDDS_OK 0
#define PDDS_ERR_INVALID 1      /* bad argument or descriptor */
#define PDDS_ERR_NOMEM 2        /* device or host allocation failed */
#define PDDS_ERR_DEVICE 3       /* HIP runtime error */
#define PDDS_ERR_UNSUPPORTED 4  /* valid request the engine does not implement */
#define PDDS_ERR_SINGULAR 5     /* an RBF interpolation matrix Psi is not invertible */
#define PDDS_ERR_NOT_READY 6    /* serving: ticket not finished yet */
#define PDDS_ERR_FULL 7         /* serving: queue full / request exceeds the HBM budget */

typedef struct pdds_engine pdds_engine;
typedef struct pdds_mesh pdds_mesh;
typedef struct pdds_lut pdds_lut;

/* ---- lifecycle ---- */
pdds_status pdds_engine_create(int device, pdds_engine **out);
void pdds_engine_destroy(pdds_engine *e);
const char *pdds_last_error(const pdds_engine *e);
const char *pdds_status_string(pdds_status s);
/* number of compute units and the device name, for launch sizing and reports */
pdds_status pdds_engine_info(const pdds_engine *e, int32_t *n_cu, char *name, int32_t name_len);

/* ---- problem descriptor (reference convention 1/2 s s^T : D2u + b.Du + c u + f = 0, c <= 0) ---- */
#define PDDS_PDE_MONEGROS 0      /* BVPs/EDPs/Monegros_Poisson_CPP.hpp:12-88 */
#define PDDS_PDE_SINSIN 1        /* u = sin(pi x) sin(pi y) + x^2 + y^2, sigma = sqrt2 I, c = 0 */
#define PDDS_PDE_VARCOEF 2       /* div form 1/2 div(s s^T grad u) + b.grad u - c u = f, u = e^x cos y */
#define PDDS_PDE_VARCOEF_LUT 3   /* PDE 2 with sigma11, sigma22, b1, b2, c_ref from a 5-field LUT */
#define PDDS_PDE_HELMHOLTZ 4     /* u as PDE 1, c = -k2 */
#define PDDS_PDE_TABLE 5         /* every coefficient tabulated: lut holds sigma11, sigma22, b1, b2, c, f,
                                    du/dx, du/dy (8 fields; sigma diagonal) and ulut the solution u, the
                                    Dirichlet data g (1 field): the texture-served problem of the
                                    reference's MCinCUDA texMode 4 (Solvers/MCinCUDA.cu:453-478), which
                                    serves f, c and grad u from textures, with sigma and b tabulated too */
/* Boundary sides of the global rectangle, the bits of pdds_problem.refl */
#define PDDS_SIDE_N 1
#define PDDS_SIDE_S 2
#define PDDS_SIDE_E 4
#define PDDS_SIDE_W 8
typedef struct {
    int32_t pde;
    double k2;              /* PDDS_PDE_HELMHOLTZ only */
    const pdds_lut *lut;    /* PDDS_PDE_VARCOEF_LUT only */
    /* Reflecting (Neumann / Robin) sides, the reference's bvp.Neumann (BVPs/BVP.hpp:133) with
       psi and varphi (:125-130): paths reflect off the global sides in refl (PDDS_SIDE_* bits;
       0: every side Dirichlet; one side at least must absorb, so refl != 15) by the Lepingle step
       (Integrators/FKACIntegrator.hpp:114-165) and stop only on the others.  There du/dn = varphi u + psi with varphi = robin (<= 0; 0 for
       Neumann) and psi taken from the problem's solution (psi = grad u . n - varphi u).  rho: the
       Lepingle step is taken within this distance of a reflecting side (the reference's rho),
       0: on every step whose nearest side reflects.  pdds_paths reads refl over its rect's own
       sides.  Rectangular domains only; knots on a reflecting side are not supported in an
       assembly. */
    int32_t refl;
    double robin;
    double rho;
    const pdds_lut *ulut;   /* PDDS_PDE_TABLE: u (1 field) */
    const pdds_lut *vlut;   /* PDDS_STEP_VRCV: mu1, mu2, F1, F2 (4 fields) */
} pdds_problem;

/* ---- run parameters ---- */
/* Step families of the Euler-Maruyama scheme (Integrators/FKACIntegrator.hpp), with mu = -sigma^T
   grad u / u from the problem's u and grad u: */
#define PDDS_STEP_CV 0          /* control variate: xi += Y (-sigma^T grad u) . dW
                                   (FKACIntegrator.hpp:46-57), the default */
#define PDDS_STEP_PLAIN 1       /* plain step, xi = 0 (FKACIntegrator.hpp:20-29): problems without a
                                   known gradient */
#define PDDS_STEP_VR 2          /* variance reduction by a change of drift: Y += Y mu . dW and
                                   X += (b - sigma mu) h + sigma dW, xi = 0 (FKACIntegrator.hpp:31-44);
                                   u must not vanish along the paths */
#define PDDS_STEP_VRCV 3        /* variance reduction and control variate with mu and F from the
                                   problem's vlut: xi += Y F . dW, Y += c Y h + Y mu . dW and
                                   X += (b - sigma mu) h + sigma dW (FKACIntegrator.hpp:100-112); any
                                   bounded mu and F leave E[score] unchanged, and F = -sigma^T grad u
                                   - u mu makes the score's martingale part vanish */
#define PDDS_FP64 0             /* FP64 path state; bit-exact with the FP64 oracle */
#define PDDS_FP32 1             /* FP32 path state, Kahan functionals, FP64 cross-path sums */
typedef struct {
    double h;               /* Euler-Maruyama step */
    double beta;            /* Gobet constant, reference -0.5826 (Solvers/EMFKACSolver.hpp:186) */
    uint32_t seed;          /* Philox key word 0 */
    uint32_t stream;        /* Philox key word 1: request id, 0 for plain assembly */
    int32_t precision;      /* PDDS_FP64 or PDDS_FP32 */
    int32_t quirk_last_side;/* 1: corner knots keep the last side's value (Meshes/stencil.cpp:376-395);
                               rectangular meshes only, disc meshes always sum */
    int64_t max_steps;      /* per-path step cap (0: 2^26) */
    int32_t sub_batch;      /* paths per variance batch S (0: 32) */
    int32_t item_paths;     /* paths per work item P, multiple of S (0: chosen from the path count) */
    int32_t wg_size;        /* work-group size (0: 256) */
    int32_t grid_wgs;       /* persistent work-groups (0: 2 per CU) */
    int32_t lut_lds;        /* 1: stage LUT tiles in LDS when they fit (default 1) */
    int32_t plain;          /* 1: plain launch -- one path per lane, one work-group per item of wg_size
                               paths, no lane recycling, no persistent queue (the baseline GPU form) */
    int32_t paths_per_lane; /* persistent launches: a wave takes 64 x this many paths from its queue
                               at once (0: 4); results do not depend on it */
    int32_t step;           /* step family (PDDS_STEP_*, default 0) */
    int32_t adaptive_h;     /* 1: each knot's paths step with h_k = min(h, (d / (2 |beta| |N^T sigma(X0)|))^2),
                               d the knot's signed distance to its stopping domain and N that side's
                               normal, when d <= -1e-4 (else h): the Gobet shift is at most half the
                               distance to the nearest side (PDDAlgorithms/PDDSparseGM.cpp:4025-4029,
                               :4386-4389).  0: every path steps with h (default).  Not with
                               reflecting sides. */
    int32_t lut_pad;        /* LDS tile shape: spare texels per tile row (0..15, default 0), which
                               move the LDS banks each row starts on; results do not depend on it */
} pdds_run;
void pdds_run_defaults(pdds_run *r);

/* ---- coefficient LUTs (bicubic B-spline, FP32 coefficients, device resident) ----
   samples: [ny][nx][nf] FP64 on the box [x0,x1]x[y0,y1].
   mode 0: node-aligned samples, interpolating spline (prefiltered coefficients);
   mode 1: texel-centred samples used raw as coefficients (cubicTex2DSimple semantics). */
pdds_status pdds_lut_create(pdds_engine *e, int32_t nx, int32_t ny, int32_t nf, const double *samples,
                            double x0, double y0, double x1, double y1, int32_t mode, pdds_lut **out);
void pdds_lut_destroy(pdds_lut *l);
/* evaluate field f of the LUT at n points on the device (FP64 arithmetic) */
pdds_status pdds_lut_eval(pdds_engine *e, const pdds_lut *l, int32_t field, int64_t n, const double *x,
                          const double *y, double *out);

/* ---- mesh: interfaces, knots, stencils, CSR pattern, Psi^-1 ---- */
pdds_status pdds_mesh_create(double sw_x, double sw_y, double ne_x, double ne_y, int32_t n_sub,
                             int32_t n_node, const double fac[4], pdds_mesh **out);
/* Disc mesh (CircleMesh layout; the reference's Meshes/Circle*.hpp are empty, so the layout is
   this engine's): a disc of `radius` at the origin cut into n_rings rings (radii radius*i/n_rings)
   x n_sectors sectors; every arc and ray-segment interface carries kmin..kmax equispaced interior
   knots (count from a hash of seed and the interface), with a knot at every crossing of an inner
   circle and a ray and a boundary knot where each ray meets the rim.  A knot's stencil is the
   annular sector of the subdomains around its interface, with sides outer arc (0), inner arc (1),
   end ray (2) and start ray (3).  The sides join interfaces of different knot counts, so each
   centre has its own Gaussian width, shape x 3.5 x its local knot spacing, all of a side's widths
   shrinking together until the side's Psi has cond below 1e11. */
pdds_status pdds_mesh_create_disc(double radius, int32_t n_rings, int32_t n_sectors, uint32_t seed, int32_t kmin,
                                  int32_t kmax, double shape, pdds_mesh **out);
/* stencil stopping domain of a knot: *geom 0 rectangle (domain[0..3] = x0, y0, x1, y1) or 1 annular
   sector (domain = r0, r1, cos a, sin a, cos b, sin b); zeros for a knot without a stencil */
pdds_status pdds_mesh_geometry(const pdds_mesh *m, int32_t knot, int32_t *geom, double domain[6]);
/* Gaussian RBF shape c of each stencil side of a knot (Compute_ipsi's fac[kind] on rectangular
   meshes, PDDAlgorithms/PDDSparseGM.cpp:816) */
pdds_status pdds_mesh_shape(const pdds_mesh *m, int32_t knot, double c[4]);
/* Gaussian width of every stencil entry of a knot, sides concatenated as in pdds_mesh_stencil
   (disc meshes: one width per centre from the local knot spacing; rectangles: the side's c) */
pdds_status pdds_mesh_entry_shapes(const pdds_mesh *m, int32_t knot, int32_t cap, double *c, int32_t *n_entries);
void pdds_mesh_destroy(pdds_mesh *m);
int32_t pdds_mesh_n_knots(const pdds_mesh *m);
int32_t pdds_mesh_n_stencils(const pdds_mesh *m);      /* distinct stencils */
/* start position (dispatch order), reported position (Node_Position), boundary flag, stencil id */
pdds_status pdds_mesh_knots(const pdds_mesh *m, double *x0, double *y0, double *px, double *py,
                            int32_t *on_boundary, int32_t *stencil_id);
/* stencil of knot k: rect[4], side sizes msz[4] and kinds[4] (order N,S,E,W); side lists
   concatenated into idx/px/py (capacity cap); *n_entries = sum msz */
pdds_status pdds_mesh_stencil(const pdds_mesh *m, int32_t knot, double rect[4], int32_t msz[4],
                              int32_t kind[4], int32_t cap, int32_t *idx, double *px, double *py,
                              int32_t *n_entries);
/* condition numbers |Psi|_F |Psi^-1|_F per side of knot k (0 for absent sides) */
pdds_status pdds_mesh_cond(const pdds_mesh *m, int32_t knot, double cond[4]);
/* Psi^-1 of side `side` (0 N, 1 S, 2 E, 3 W) of knot k, row-major m x m into out (capacity cap) */
pdds_status pdds_mesh_ipsi(const pdds_mesh *m, int32_t knot, int32_t side, int64_t cap, double *out, int32_t *msz);
/* CSR pattern of G without the identity: row_ptr[n_knots+1]; cols may be NULL to query nnz */
pdds_status pdds_mesh_csr(const pdds_mesh *m, int64_t *row_ptr, int32_t *cols, int64_t *nnz);

/* ---- counters ---- */
typedef struct {
    int64_t path_steps;     /* phase-1 Euler-Maruyama steps over all paths */
    int64_t paths;
    int64_t exits[6];       /* N, S, E, W, global boundary, unclassified */
    int64_t capped;         /* paths stopped by max_steps */
    double kernel_ms;       /* device time of the path kernel (HIP events) */
    double reduce_ms;       /* device time of the finalize kernel */
    double total_ms;        /* host wall time of the call */
    int64_t work_items;
    int32_t grid_wgs, wg_size, item_paths, sub_batch;
} pdds_counters;

/* ---- per-path records (MCinCUDA shape) ----
   Paths path0..path0+n-1 of knot id `knot` start at (x0,y0) and stop on rect[4].
   Outputs (any may be NULL): exit projection xt/yt, Y, Z, xi, steps; phase-2 outputs when
   phase2 != 0 (continuation to the true exit with beta = 0). */
pdds_status pdds_paths(pdds_engine *e, const pdds_problem *p, const pdds_run *r, double x0, double y0,
                       const double rect[4], uint32_t knot, int64_t path0, int64_t n, int32_t phase2,
                       double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps,
                       double *xt2, double *yt2, double *Y2, double *Z2, double *xi2, int64_t *steps2,
                       pdds_counters *ctr);

/* ---- one knot: G row (mesh CSR order), G variance, B, BB ---- */
pdds_status pdds_solve_knot(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                            int32_t knot, int64_t n_paths, double *G, double *G_var, double *B, double *BB,
                            pdds_counters *ctr);

/* ---- the step each knot's paths take under r (r->h, or the adaptive_h rule), h[n_knots], 0 for
   boundary knots: the h column of the reference's Compute_h_N output
   (PDDAlgorithms/PDDSparseGM.cpp:4386-4390) ---- */
pdds_status pdds_knot_steps(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r, double *h);

/* ---- every knot of a mesh: G values in the mesh CSR pattern, per-entry per-path variance,
   B and B second moment per knot.  paths_per_knot may be NULL (n_paths each, n_paths >= 1).
   A count of 0 for a Monte Carlo knot leaves it out of this call (its row, variances, B and BB
   are written as 0): one shard of a split by knots.  A knot's results do not depend on which
   other knots share the call, so the shards' rows put together equal one call over all knots.
   ---- */
pdds_status pdds_assemble(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                          const int64_t *paths_per_knot, int64_t n_paths, double *G, double *G_var,
                          double *B, double *BB, pdds_counters *ctr);

/* ---- assembly plus each knot's equation statistics ----
   As pdds_assemble.  With u_ref (n_knots interface values, e.g. a previous solve) it also returns,
   per Monte Carlo knot k, the mean (eq_mean, may be NULL) and per-path variance (eq_var) of the
   knot's equation estimator B_k - sum_j G_kj u_ref_j, by batch means over the sub-batches;
   boundary knots get g and 0.  sqrt(eq_var / N_k) is the standard error of knot k's equation.
   Replaces the variance inputs of the commented Compute_h_N
   (PDDAlgorithms/PDDSparseGM.cpp:4338-4395). */
pdds_status pdds_assemble_eq(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                             const int64_t *paths_per_knot, int64_t n_paths, const double *u_ref, double *G,
                             double *G_var, double *B, double *BB, double *eq_mean, double *eq_var,
                             pdds_counters *ctr);

/* ---- variance-targeted path counts (batching form; the reference's policy is commented out,
   PDDAlgorithms/PDDSparseGM.cpp:4358-4365) ----
   A pilot assembly, then rounds that move each knot's path count to growth * var_k / se_target^2
   (within [n_min, n_max]) until every knot's standard error sqrt(var_k / N_k) is at most se_target
   or its count is n_max.  var_k is the batch-means variance of
     variance = 0 (default): the knot's score, control variate included -- the estimator variance
                 Var(phi) (1 - rho^2) of the reference's rule (PDDSparseGM.cpp:4358-4359);
     variance = 1: the knot's equation B_k - sum_j G_kj u_ref_j, u_ref the path-weighted mean of
                 the rounds' solutions (a pilot solve first).  Far smaller than the score's where
                 u_ref is accurate, but the stencil sides' RBF cardinal functions amplify the
                 knot-to-knot noise of a u_ref solved from few paths per knot, so the measured
                 variance can exceed the true one many times over (seen on disc meshes).  Every round is one assembly of all knots with
   their uneven counts in one persistent launch, followed by the device solve.  Round a (0 = the
   pilot) draws its paths from Philox stream r->stream + a * 0x9E3779B9 (mod 2^32), so the counts
   of a round, set from earlier rounds' paths, never depend on its own paths (re-using the pilot's
   paths would bias the estimate: a knot whose first paths happen to vary more gets more paths
   added to them).  Outputs: u (n_knots), the final counts (0 for boundary knots) and, if se is not
   NULL, each knot's standard error; u is the solve of one pdds_assemble with those counts and
   rep->stream.  A knot's standard error is measured on the final round's own paths, so it
   scatters around the value its count was set for by the sampling error of the batch-means
   variance (about sqrt(2 / batches)); growth is the margin against it, and rep->knots_over counts
   the knots that end above se_target. */
typedef struct {
    double se_target;       /* target standard error of every knot equation (default 1e-3) */
    int64_t n_min, n_max;   /* per-knot path-count bounds (defaults 1e3, 1e6) */
    int64_t n_pilot;        /* pilot paths per knot (0: n_min) */
    int32_t max_rounds;     /* refinement rounds after the pilot (default 3) */
    double tol;             /* BiCGSTAB relative tolerance (default 1e-10) */
    double growth;          /* safety factor on the variance-targeted count (default 1.1) */
    int32_t variance;       /* 0: score variance (default), 1: equation variance */
} pdds_adaptive;
typedef struct {
    int32_t rounds;         /* refinement rounds run */
    int32_t iters;          /* BiCGSTAB iterations of the final solve */
    double relres;          /* and its relative residual */
    int64_t paths;          /* paths of the final assembly */
    int64_t path_steps;     /* path-steps of every assembly, pilots included */
    double max_se;          /* largest knot standard error of the final assembly */
    int32_t knots_capped;   /* knots held at n_max above the target */
    double kernel_ms;       /* device time of the path and reduction kernels, all assemblies */
    double total_ms;        /* host wall time */
    uint32_t stream;        /* Philox stream of the final assembly */
    int32_t knots_over;     /* knots whose final standard error exceeds se_target */
} pdds_adaptive_report;
void pdds_adaptive_defaults(pdds_adaptive *c);
pdds_status pdds_solve_adaptive(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                                const pdds_adaptive *cfg, double *u, int64_t *paths_per_knot, double *se,
                                pdds_adaptive_report *rep);

/* ---- launch autotuner (no counterpart in the reference) ----
   Times trial assemblies of a sample of the mesh's knots (a contiguous run of at most max_knots
   Monte Carlo knots from the middle of the mesh, sample_paths paths each, so that the sample keeps
   the mesh's stencil and LUT-tile structure) over the launch shapes that do not change results
   -- LDS tiling on or off, work-group size, work-groups per CU, paths per lane -- and writes into
   *best the base run with the fastest shape (path-steps per second of the whole assembly). */
typedef struct {
    int64_t sample_paths;   /* paths per sampled knot (0: 16384) */
    int32_t max_knots;      /* sampled knots (0: 512) */
    int32_t reps;           /* timed repetitions per candidate, after one warm-up (0: 2) */
} pdds_tune_config;
typedef struct {
    int32_t candidates;     /* shapes timed */
    double best_rate;       /* path-steps/s of the chosen shape */
    double base_rate;       /* path-steps/s of the base run's own shape */
    double total_ms;        /* host wall time of the search */
} pdds_tune_report;
pdds_status pdds_autotune(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *base,
                          const pdds_tune_config *cfg, pdds_run *best, pdds_tune_report *rep);

/* ---- Feynman-Kac estimate of u at n points of the global rectangle `glob` ---- */
pdds_status pdds_point_eval(pdds_engine *e, const pdds_problem *p, const pdds_run *r, const double glob[4],
                            int32_t n, const double *x, const double *y, const int64_t *n_paths,
                            double *u, double *u_var, pdds_counters *ctr);

/* ---- point evaluation that allocates no device memory ----
   No counterpart in the reference.  For callers that cannot allocate device memory: code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Arguments, results
   and errors are those of pdds_point_eval (bit-identical results), but every device buffer the
   call needs (exit records, work items, reduction partials, statistics, queue counters and the
   rectangle's stencil table) is taken from one fixed array of PDDS_NOALLOC_SCRATCH_BYTES that
   each device holds from the library's load until the process ends; the call makes no hipMalloc
   or hipFree.  It returns PDDS_ERR_FULL, before any device work, when its buffers would exceed
   that array: about 32 bytes per path (some 2e6 paths per call in total).  Every call on one
   device shares the array, which keeps the last call's contents (what must start at zero is
   cleared by each call); overlapping calls on one device would race, so the library serialises
   them with a per-device lock (they wait for each other, also across engines). */
#define PDDS_NOALLOC_SCRATCH_BYTES ((int64_t)64 << 20)
pdds_status pdds_point_eval_noalloc(pdds_engine *e, const pdds_problem *p, const pdds_run *r,
                                    const double glob[4], int32_t n, const double *x, const double *y,
                                    const int64_t *n_paths, double *u, double *u_var, pdds_counters *ctr);

/* ---- (I + G) u = B by BiCGSTAB on the device, zero initial guess ----
   rows given without the identity (the mesh CSR pattern). Returns iterations in *iters. */
pdds_status pdds_solve(pdds_engine *e, int32_t n, const int64_t *row_ptr, const int32_t *cols,
                       const double *G, const double *B, double *u, double tol, int32_t maxit,
                       int32_t *iters, double *relres);

/* ---- serving (no counterpart in the reference): a continuous-batching request loop ----
   Requests are point evaluations of u or full interface solves on the rectangle-mesh
   geometry.  A host pool prepares each request (a solve builds its mesh and Psi^-1 tables);
   one device thread then runs rounds: each round admits prepared requests in arrival order
   while their device footprint fits hbm_budget and fewer than max_inflight are admitted,
   packs the not-yet-run work of every admitted request (point paths first, then solve knots,
   at most round_paths paths) into one persistent launch, and completes every point whose paths
   are done; a solve whose knots are all done goes to a solver thread, which runs its device
   BiCGSTAB on a second engine (its own stream, default priority) while the next round's paths
   run.  A solve larger than a round is spread over rounds, so requests arriving meanwhile join
   the next round.  The paths of a request
   use its own Philox stream, and each of its sums has a fixed order: its result is the same,
   bit for bit, alone or in any mix.  The server uses the engine's stream and lock; other
   engine calls made meanwhile wait for the current round. */
typedef struct pdds_server pdds_server;
typedef struct {
    pdds_problem problem;   /* PDE of every request (the Helmholtz k2 comes with each request) */
    pdds_run run;           /* h, seed, precision and launch parameters; stream is per request */
    double domain[4];       /* x0, y0, x1, y1 of every request's problem */
    int64_t hbm_budget;     /* device bytes for admitted requests (0: 16 GiB): each is counted by its
                               stencil tables, the exit records of at most one round of its paths,
                               its reduction partials and descriptors, and its G values -- an
                               estimate of what it adds to the engine's buffers */
    int32_t max_inflight;   /* requests admitted at once (0: 256) */
    int32_t prep_threads;   /* host threads preparing requests (0: 2); a solve's mesh build already
                               spreads over up to 16 host threads, so more workers mostly share
                               the same cores and delay every request alike */
    int64_t round_paths;    /* paths per round (0: 2^25) */
} pdds_server_config;

#define PDDS_REQ_POINT 0
#define PDDS_REQ_SOLVE 1
typedef struct {
    int32_t kind;           /* PDDS_REQ_POINT or PDDS_REQ_SOLVE */
    uint32_t stream;        /* Philox key word 1 of this request's paths */
    double k2;              /* Helmholtz parameter (PDDS_PDE_HELMHOLTZ) */
    int64_t n_paths;        /* paths (point) or paths per Monte Carlo knot (solve) */
    double x, y;            /* point */
    int32_t n_sub, n_node;  /* solve: subdomains per side, knots per interface */
    double fac[4];          /* solve: RBF shape per stencil kind */
    double tol;             /* solve: BiCGSTAB relative tolerance (0: 1e-10) */
    double h;               /* Euler-Maruyama step of this request's paths (0: the server run's h); the
                               step bias of a solve grows as its subdomains shrink, so a fixed error
                               target asks for h in proportion to the squared subdomain width */
} pdds_request;

typedef struct {
    int32_t kind;
    pdds_status status;     /* PDDS_OK, or why the request failed (a singular mesh, ...) */
    int32_t n_knots;        /* solve: unknowns (u holds min(n_knots, cap) of them, knot order) */
    int32_t iters;          /* solve: BiCGSTAB iterations */
    double relres;          /* solve: final relative residual */
    double u, u_var;        /* point: estimate and per-path variance */
    int64_t path_steps;
    double queue_ms;        /* submit -> admitted to a round */
    double latency_ms;      /* submit -> done */
} pdds_result;

pdds_status pdds_server_create(pdds_engine *e, const pdds_server_config *cfg, pdds_server **out);
/* finishes the requests already submitted, then stops */
void pdds_server_destroy(pdds_server *s);
/* PDDS_ERR_FULL if the request alone exceeds the HBM budget */
pdds_status pdds_server_submit(pdds_server *s, const pdds_request *r, uint64_t *ticket);
/* PDDS_ERR_NOT_READY while the request runs; once done fills *res (and u for a solve) and
   releases the ticket */
pdds_status pdds_server_poll(pdds_server *s, uint64_t ticket, pdds_result *res, double *u, int64_t cap);
/* as poll, waiting for the request; PDDS_ERR_INVALID if another thread took the result meanwhile */
pdds_status pdds_server_wait(pdds_server *s, uint64_t ticket, pdds_result *res, double *u, int64_t cap);
/* totals since creation: requests submitted and done, launch rounds, path-steps, device busy ms */
pdds_status pdds_server_stats(pdds_server *s, int64_t *submitted, int64_t *done, int64_t *rounds,
                              int64_t *path_steps, double *busy_ms);

#ifdef __cplusplus
}
#endif
#endif
