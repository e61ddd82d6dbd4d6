/*
 * oracle.c -- TEST INFRASTRUCTURE (the oracle). Not part of the product.
 *
 * Plain-C, OpenMP, FP64 restatement of the reference's PDDSparse CPU path.
 * See oracle.h for the list of functions and how each one is pinned.
 */
#include "oracle.h"
#include "orc_math.h"
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* elementary pieces                                                         */
/* ------------------------------------------------------------------------ */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    orc_philox4x32_10(ctr, key, out);
}

void orc_math_eval(int fn, const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: y[i] = orc_exp(x[i]); break;
            case 1: y[i] = orc_log(x[i]); break;
            case 2: y[i] = orc_sin(x[i]); break;
            case 3: y[i] = orc_cos(x[i]); break;
            case 4: y[i] = orc_tanh(x[i]); break;
            case 5: y[i] = orc_sinpi(x[i]); break;
            case 6: y[i] = orc_cospi(x[i]); break;
            case 8: y[i] = orc_exp_tab(x[i]); break;
            default: y[i] = orc_expm1(x[i]); break;
        }
    }
}

/* Normal pair for step `step` of path `path` of knot `knot`.
   Keying: key = (seed, stream), counter = (call, path, knot, 0).
   normals == 0: call = step, both normals from one 4-word draw (53-bit uniforms).
   normals == 1: call = step/2, words (0,1) feed even steps, (2,3) odd steps (24-bit uniforms).
   Replaces Increment_Update (Integrators/FKACIntegrator.hpp:15-19). */
static inline void orc_normal_pair(const orc_run *run, uint32_t knot, uint32_t path, int64_t step,
                                   double *n1, double *n2) {
    uint32_t key[2] = {run->seed, run->stream};
    uint32_t w[4];
    if (run->normals == 0) {
        uint32_t ctr[4] = {(uint32_t)step, path, knot, (uint32_t)(step >> 32)};
        orc_philox4x32_10(ctr, key, w);
        orc_box_muller(orc_u01_53(w[0], w[1]), orc_u01_53(w[2], w[3]), n1, n2);
    } else {
        int64_t call = step >> 1;
        uint32_t ctr[4] = {(uint32_t)call, path, knot, (uint32_t)(call >> 32)};
        orc_philox4x32_10(ctr, key, w);
        if ((step & 1) == 0) orc_box_muller(orc_u01_24(w[0]), orc_u01_24(w[1]), n1, n2);
        else orc_box_muller(orc_u01_24(w[2]), orc_u01_24(w[3]), n1, n2);
    }
}

void orc_normals(const orc_run *run, uint32_t knot, uint32_t path, int64_t step0, int64_t nsteps,
                 double *n1, double *n2) {
    for (int64_t s = 0; s < nsteps; ++s) orc_normal_pair(run, knot, path, step0 + s, &n1[s], &n2[s]);
}

/* Signed distance to the rectangle p = [x0,y0,x1,y1], nearest-side projection and
   outward normal.  Restates Rectangle2D (BVPs/Domains/rectangle.hpp:7-229) with the
   same operation order (centre/half-side form, quadrant split, inside tie rule
   "dx-hx > dy-hy", corner normal 0.70710678118f). */
double orc_rect_dist(const double p[4], double x, double y, double *ex, double *ey, double *nx, double *ny) {
    const double cx = 0.5 * (p[0] + p[2]), cy = 0.5 * (p[1] + p[3]);
    const double hx = p[2] - cx, hy = p[3] - cy;
    const double dx = fabs(x - cx), dy = fabs(y - cy);
    const double diag = (double)0.70710678118f;
    const double sxs = (x > cx) ? 1.0 : -1.0;  /* right half: quadrants 1,4 */
    const double sys = (y > cy) ? 1.0 : -1.0;  /* top half: quadrants 1,2 */
    *ex = x; *ey = y;
    if (dx <= hx && dy <= hy) {
        if (dx - hx > dy - hy) {
            *ex = sxs * hx + cx; *nx = sxs; *ny = 0.0;
            return dx - hx;
        }
        *ey = sys * hy + cy; *nx = 0.0; *ny = sys;
        return dy - hy;
    }
    if (dx > hx) {
        if (dy > hy) {
            double r = sqrt((dx - hx) * (dx - hx) + (dy - hy) * (dy - hy));
            *ex = sxs * hx + cx; *ey = sys * hy + cy;
            *nx = sxs * diag; *ny = sys * diag;
            return r;
        }
        *ex = sxs * hx + cx; *nx = sxs; *ny = 0.0;
        return dx - hx;
    }
    *ey = sys * hy + cy; *nx = 0.0; *ny = sys;
    return dy - hy;
}

/* Signed distance to the annular sector g = {r0, r1, cos a, sin a, cos b, sin b} of a disc mesh
   (the engine's own CircleMesh layout; the reference's Meshes/Circle*.hpp files are empty):
   the largest of the four constraint violations r - r1, r0 - r (none when r0 = 0), and the
   signed distances beyond the rays a and b.  Projection onto that side, clamped to it, and its
   outward normal; kind 0 outer arc, 1 inner arc, 2 ray b, 3 ray a. */
double orc_sector_dist(const double g[6], double x, double y, double *ex, double *ey, double *nx, double *ny,
                       int *kind) {
    const double r0 = g[0], r1 = g[1], ca = g[2], sa = g[3], cb = g[4], sb = g[5];
    const double r = sqrt(x * x + y * y);
    const double sA = ca * y - sa * x, sB = x * sb - y * cb;
    const int whole = r0 < 0.0;   /* the whole disc r <= r1: its circle is the only side */
    double d = r - r1;
    int k = 0;
    const double din = r0 > 0.0 ? r0 - r : -(r + r1);
    if (din > d) { d = din; k = 1; }
    if (!whole && -sB > d) { d = -sB; k = 2; }
    if (!whole && -sA > d) { d = -sA; k = 3; }
    if (k <= 1) {
        double ux = x / r, uy = y / r;
        if (!whole && sA < 0.0) { ux = ca; uy = sa; }
        else if (!whole && sB < 0.0) { ux = cb; uy = sb; }
        const double rr = k == 0 ? r1 : r0;
        *ex = rr * ux; *ey = rr * uy;
        *nx = k == 0 ? ux : -ux; *ny = k == 0 ? uy : -uy;
    } else {
        const double dx = k == 3 ? ca : cb, dy = k == 3 ? sa : sb;
        double t = x * dx + y * dy;
        t = t < r0 ? r0 : (t > r1 ? r1 : t);
        *ex = t * dx; *ey = t * dy;
        *nx = k == 3 ? sa : -sb;
        *ny = k == 3 ? -ca : cb;
    }
    if (kind) *kind = k;
    return d;
}

/* distance to a stencil's stopping domain: geom 0 rectangle dom[4], geom 1 sector dom[6] */
static inline double orc_domain_dist(int geom, const double *dom, double x, double y, double *ex, double *ey,
                                     double *nx, double *ny) {
    return geom ? orc_sector_dist(dom, x, y, ex, ey, nx, ny, NULL) : orc_rect_dist(dom, x, y, ex, ey, nx, ny);
}

/* ------------------------------------------------------------------------ */
/* PDE coefficients (reference convention, BVPs/BVP.hpp:89-137)              */
/* ------------------------------------------------------------------------ */
#define MK1 1.0
#define MK2 0.01
#define MK3 0.02
#define MK4 0.12
#define MK5 0.05
#define MK6 0.05
#define MK7 (-0.12)
#define MSF 0.333

/* Monegros Poisson: u, grad u, Laplacian (BVPs/EDPs/Monegros_Poisson_CPP.hpp:12-40, 57-59) */
static void orc_monegros(double x, double y, double *u, double *gx, double *gy, double *lap) {
    double a1 = sqrt(MK1 + MK2 * x * x + MK3 * y * y);
    double s1 = orc_sin(MK4 * x + MK5 * y), s2 = orc_sin(MK6 * x + MK7 * y);
    double c1 = orc_cos(MK4 * x + MK5 * y), c2 = orc_cos(MK6 * x + MK7 * y);
    double T = orc_tanh(s1 + s2);
    double sa = orc_sin(a1), ca = orc_cos(a1);
    double T21 = T * T - 1.0;
    double px = MK4 * c1 + MK6 * c2, py = MK5 * c1 + MK7 * c2;
    double a3 = a1 * a1 * a1;
    *u = 3.0 + MSF * (sa + T);
    *gx = MSF * ((ca * MK2 * x / a1) - px * T21);
    *gy = MSF * (ca * MK1 * MK3 * y / a1 - py * T21);
    double dxx = MSF * (((MK2 / a1) - (MK2 * MK2 * x * x / a3)) * ca - (x * x * MK2 * MK2 / (a1 * a1)) * sa
                        + 2.0 * px * px * T21 * T + (MK4 * MK4 * s1 + MK6 * MK6 * s2) * T21);
    double dyy = MSF * (((MK3 / a1) - (MK3 * MK3 * y * y / a3)) * ca - (y * y * MK3 * MK3 / (a1 * a1)) * sa
                        + 2.0 * py * py * T21 * T + (MK5 * MK5 * s1 + MK7 * MK7 * s2) * T21);
    *lap = dxx + dyy;
}

/* Uniform cubic B-spline weights of taps -1..2 (Horner forms with fmaf) */
static void orc_bspline_wf(float f, float w[4]) {
    const float s6 = 1.0f / 6.0f;
    const float f2 = f * f, om = 1.0f - f;
    w[0] = (om * om * om) * s6;
    w[1] = fmaf(fmaf(3.0f, f, -6.0f), f2, 4.0f) * s6;
    w[2] = fmaf(fmaf(fmaf(-3.0f, f, 3.0f), f, 3.0f), f, 1.0f) * s6;
    w[3] = (f2 * f) * s6;
}

/* Bicubic B-spline lookup: the coordinate in FP64, the 16-tap weighted sum of the FP32
   coefficients in FP32 with fma -- the arithmetic of the reference's FP32 texture path
   (Integrators/cubicTex2D.cu:64-87, Integrators/bspline_kernel.cu:52-60).  lut_nf is the
   texel stride in floats. */
static double orc_lut_at(const orc_lut *L, int field, double x, double y) {
    const int nx = L->nx, ny = L->ny, nf = L->nf;
    double t = (x - L->ox) * L->sx;
    double v = (y - L->oy) * L->sy;
    t = fmin(fmax(t, -2.0), (double)nx + 1.0);
    v = fmin(fmax(v, -2.0), (double)ny + 1.0);
    double fi = floor(t), fj = floor(v);
    int i = (int)fi, j = (int)fj;
    float wx[4], wy[4];
    orc_bspline_wf((float)(t - fi), wx);
    orc_bspline_wf((float)(v - fj), wy);
    float acc = 0.0f;
    for (int b = 0; b < 4; ++b) {
        int jj = j - 1 + b; jj = jj < 0 ? 0 : (jj > ny - 1 ? ny - 1 : jj);
        float row = 0.0f;
        for (int a = 0; a < 4; ++a) {
            int ii = i - 1 + a; ii = ii < 0 ? 0 : (ii > nx - 1 ? nx - 1 : ii);
            row = fmaf(wx[a], L->c[((int64_t)jj * nx + ii) * nf + field], row);
        }
        acc = fmaf(wy[b], row, acc);
    }
    return (double)acc;
}

double orc_lut_eval(const orc_problem *pb, int field, double x, double y) {
    const orc_lut L = {pb->lut, pb->lut_nx, pb->lut_ny, pb->lut_nf, pb->lut_ox, pb->lut_oy, pb->lut_sx, pb->lut_sy};
    return orc_lut_at(&L, field, x, y);
}

/* Coefficients at (x,y): sigma (row-major s00 s01 s10 s11), b, c, f, grad u, g, u */
void orc_pde_eval(const orc_problem *pb, double x, double y, double sig[4], double b[2],
                  double *c, double *f, double grad[2], double *g, double *u) {
    switch (pb->pde) {
        case ORC_PDE_MONEGROS: {
            double uu, gx, gy, lap;
            orc_monegros(x, y, &uu, &gx, &gy, &lap);
            sig[0] = 1.41421356237; sig[1] = 0.0; sig[2] = 0.0; sig[3] = 1.41421356237;
            b[0] = 0.0; b[1] = 0.0; *c = -0.0; *f = -lap;
            grad[0] = gx; grad[1] = gy; *g = uu; *u = uu;
            break;
        }
        case ORC_PDE_SINSIN:
        case ORC_PDE_HELMHOLTZ: {
            double sx, cx, sy, cy;
            orc_sincospi(x, &sx, &cx);
            orc_sincospi(y, &sy, &cy);
            double uu = sx * sy + x * x + y * y;
            sig[0] = 1.4142135623730951; sig[1] = 0.0; sig[2] = 0.0; sig[3] = 1.4142135623730951;
            b[0] = 0.0; b[1] = 0.0;
            grad[0] = ORC_PI * cx * sy + 2.0 * x;
            grad[1] = ORC_PI * sx * cy + 2.0 * y;
            double flap = 19.739208802178716 * sx * sy - 4.0;  /* -Laplacian */
            if (pb->pde == ORC_PDE_SINSIN) { *c = 0.0; *f = flap; }
            else { *c = -pb->k2; *f = flap + pb->k2 * uu; }
            *g = uu; *u = uu;
            break;
        }
        case ORC_PDE_TABLE: {
            /* the texture-served problem (Solvers/MCinCUDA.cu:453-478, texMode 4: f, c and
               grad u from textures), sigma and b tabulated as well */
            sig[0] = orc_lut_eval(pb, 0, x, y); sig[1] = 0.0; sig[2] = 0.0;
            sig[3] = orc_lut_eval(pb, 1, x, y);
            b[0] = orc_lut_eval(pb, 2, x, y);
            b[1] = orc_lut_eval(pb, 3, x, y);
            *c = orc_lut_eval(pb, 4, x, y);
            *f = orc_lut_eval(pb, 5, x, y);
            grad[0] = orc_lut_eval(pb, 6, x, y);
            grad[1] = orc_lut_eval(pb, 7, x, y);
            *g = orc_lut_at(&pb->ulut, 0, x, y);
            *u = *g;
            break;
        }
        default: { /* VARCOEF and VARCOEF_LUT */
            double sxn = orc_sin(x), cxn = orc_cos(x), syn = orc_sin(y), cyn = orc_cos(y);
            double ex = orc_exp(x);
            double s1 = 1.0 + 0.3 * sxn, s2 = 1.0 + 0.3 * cyn;
            double uu = ex * cyn, uy = -ex * syn;
            double a1 = s1 * s1, a2 = s2 * s2;
            double a1p = 2.0 * s1 * (0.3 * cxn), a2p = 2.0 * s2 * (-0.3 * syn);
            double fcfg = 0.5 * (a1p * uu + a1 * uu + a2p * uy - a2 * uu)
                          + 0.2 * y * uu + (-0.2 * x) * uy - (1.0 + x * x) * uu;
            *f = -fcfg;
            grad[0] = uu; grad[1] = uy; *g = uu; *u = uu;
            if (pb->pde == ORC_PDE_VARCOEF) {
                sig[0] = s1; sig[1] = 0.0; sig[2] = 0.0; sig[3] = s2;
                b[0] = 0.2 * y + 0.3 * s1 * cxn;
                b[1] = -0.2 * x - 0.3 * s2 * syn;
                *c = -(1.0 + x * x);
            } else {
                sig[0] = orc_lut_eval(pb, 0, x, y); sig[1] = 0.0; sig[2] = 0.0;
                sig[3] = orc_lut_eval(pb, 1, x, y);
                b[0] = orc_lut_eval(pb, 2, x, y);
                b[1] = orc_lut_eval(pb, 3, x, y);
                *c = orc_lut_eval(pb, 4, x, y);
            }
            break;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Monte Carlo: EMFKACSolver::Simulate_OMP (Solvers/EMFKACSolver.hpp:108-216) */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, Y, Z, xi; double sig[4], b[2], c, f, grad[2], mu[2], F[2]; } orc_state;

static inline void orc_coeffs(const orc_problem *pb, orc_state *s) {
    double g, u;
    orc_pde_eval(pb, s->x, s->y, s->sig, s->b, &s->c, &s->f, s->grad, &g, &u);
    if (pb->step == 1) {
        s->grad[0] = 0.0;
        s->grad[1] = 0.0;
    } else if (pb->step == 2) {
        s->grad[0] = s->grad[0] / g;   /* g = u for every manufactured problem */
        s->grad[1] = s->grad[1] / g;
    } else if (pb->step == 3) {
        s->mu[0] = orc_lut_at(&pb->vlut, 0, s->x, s->y);
        s->mu[1] = orc_lut_at(&pb->vlut, 1, s->x, s->y);
        s->F[0] = orc_lut_at(&pb->vlut, 2, s->x, s->y);
        s->F[1] = orc_lut_at(&pb->vlut, 3, s->x, s->y);
    }
}

/* One Euler-Maruyama step: all updates use the old X and Y; Dirichlet everywhere so
   psi = varphi = 0.  Coefficients at the old X are already in s, with v = -sigma^T grad.
   Control-variate and plain steps (Integrators/FKACIntegrator.hpp:46-57, :20-29):
   xi += Y v . dW.  VR step (:31-44, grad holds grad u / u so v = mu): Y += c Y h + Y mu . dW
   and X's drift is b - sigma mu.  VR+CV step (:100-112, mu and F from the table): xi += Y F . dW,
   Y += c Y h + Y mu . dW, X's drift b - sigma mu. */
static inline void orc_step(int step, orc_state *s, double h, double sqrth, double n1, double n2, double lz,
                            double ly) {
    /* lz = psi Y L, ly = varphi Y L: the local-time terms of a reflecting side (0 elsewhere),
       Z += f Y h + psi Y L and Y += c Y h + varphi Y L (Integrators/FKACIntegrator.hpp:156-158) */
    double i0 = sqrth * n1, i1 = sqrth * n2;
    double v0 = -(s->sig[0] * s->grad[0] + s->sig[2] * s->grad[1]);
    double v1 = -(s->sig[1] * s->grad[0] + s->sig[3] * s->grad[1]);
    if (step == 3) { v0 = s->F[0]; v1 = s->F[1]; }
    double dxi = s->Y * (v0 * i0 + v1 * i1);
    if (step != 2) s->xi = s->xi + dxi;
    s->Z = s->Z + (s->f * s->Y * h + lz);
    double Yold = s->Y;
    double b0 = s->b[0], b1 = s->b[1];
    if (step == 2) {
        s->Y = Yold + ((s->c * Yold * h + dxi) + ly);
        b0 = b0 - (s->sig[0] * v0 + s->sig[1] * v1);
        b1 = b1 - (s->sig[2] * v0 + s->sig[3] * v1);
    } else if (step == 3) {
        const double m0 = s->mu[0], m1 = s->mu[1];
        s->Y = Yold + ((s->c * Yold * h + Yold * (m0 * i0 + m1 * i1)) + ly);
        b0 = b0 - (s->sig[0] * m0 + s->sig[1] * m1);
        b1 = b1 - (s->sig[2] * m0 + s->sig[3] * m1);
    } else {
        s->Y = Yold + (s->c * Yold * h + ly);
    }
    double nx = s->x + (b0 * h + (s->sig[0] * i0 + s->sig[1] * i1));
    double ny = s->y + (b1 * h + (s->sig[2] * i0 + s->sig[3] * i1));
    s->x = nx; s->y = ny;
}

/* Inside() with the Gobet shift (Integrators/FKACIntegrator.hpp:168-186): returns 1 while the
   path continues, i.e. dist < beta * |N^T sigma(X)| * sqrt(h).  Evaluates the coefficients at
   the new X (sigma is needed here, the rest by the next step). */
static inline int orc_inside(const orc_problem *pb, orc_state *s, int geom, const double *dom, double beta,
                             double sqrth, double *ex, double *ey) {
    double nx, ny;
    double d = orc_domain_dist(geom, dom, s->x, s->y, ex, ey, &nx, &ny);
    orc_coeffs(pb, s);
    double a = nx * s->sig[0] + ny * s->sig[2];
    double b = nx * s->sig[1] + ny * s->sig[3];
    double nrm = sqrt(a * a + b * b);
    return d < beta * nrm * sqrth;
}

/* ---- reflecting sides: the Lepingle step (Integrators/FKACIntegrator.hpp:114-165) ---- */
/* Nearest side of a rectangle named by its outward normal (orc_rect_dist's): 0 N, 1 S, 2 E, 3 W,
   corners 4 NE, 5 NW, 6 SE, 7 SW. */
static inline int orc_side_of_normal(double nx, double ny) {
    if (nx == 0.0) return ny > 0.0 ? 0 : 1;
    if (ny == 0.0) return nx > 0.0 ? 2 : 3;
    return ny > 0.0 ? (nx > 0.0 ? 4 : 5) : (nx > 0.0 ? 6 : 7);
}
/* whether that side reflects (a corner: both of its sides) */
static inline int orc_reflects(int kind, int mask) {
    static const int sides[8][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 2}, {0, 3}, {1, 2}, {1, 3}};
    return ((mask >> sides[kind][0]) & 1) && ((mask >> sides[kind][1]) & 1);
}
/* E ~ Exp(1) for step `step`: a Philox draw of its own, counter word 3 = 2^31 */
static inline double orc_exp_draw(const orc_run *run, uint32_t knot, uint32_t path, int64_t step) {
    uint32_t key[2] = {run->seed, run->stream};
    uint32_t ctr[4] = {(uint32_t)step, path, knot, 0x80000000u};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    return -orc_log(run->normals == 0 ? orc_u01_53(w[0], w[1]) : orc_u01_24(w[0]));
}
/* One step from X whose nearest side (outward normal n, signed distance d) reflects.  Lepingle
   (lep): the maximum of the step's Brownian bridge in the normal direction is
   M = (u + sqrt(u^2 + 2 |n^T sigma|^2 h E)) / 2, u the normal displacement of the free step;
   the local time L = max(0, M + d) pushes the new X back along n.  The functionals take L and
   the pending local time lt, with psi at the old X's projection onto the side (the reference
   takes it at the old X itself, O(sqrt h) inside, which costs the estimate half an order). */
static inline void orc_refl_step(const orc_problem *pb, orc_state *s, double h, double sqrth, double n1, double n2,
                                 int lep, double E, double nx, double ny, double d, double lt) {
    double L = 0.0;
    if (lep) {
        const double i0 = sqrth * n1, i1 = sqrth * n2;
        const double ga = nx * s->sig[0] + ny * s->sig[2], gb = nx * s->sig[1] + ny * s->sig[3];
        const double uc = (ga * i0 + gb * i1) + (nx * s->b[0] + ny * s->b[1]) * h;
        const double M = 0.5 * (uc + sqrt((ga * ga + gb * gb) * ((2.0 * h) * E) + uc * uc));
        L = fmax(0.0, M + d);
    }
    const double Lt = L + lt;
    double lz = 0.0, ly = 0.0;
    if (Lt > 0.0) {
        /* boundary data at X's projection onto the side, X - d n */
        double sig[4], b[2], c, f, grad[2], g, u;
        orc_pde_eval(pb, s->x - d * nx, s->y - d * ny, sig, b, &c, &f, grad, &g, &u);
        const double psi = (grad[0] * nx + grad[1] * ny) - pb->robin * g;
        /* the local time weighed by (e^{varphi L} - 1) / (varphi L): Y's exact growth over it
           (1 on a Neumann side; the reference's linear Y += varphi Y L costs half an order) */
        const double x = pb->robin * Lt;
        const double w = Lt * (x == 0.0 ? 1.0 : orc_expm1(x) / x);
        lz = (psi * s->Y) * w;
        ly = (pb->robin * s->Y) * w;
    }
    orc_step(pb->step, s, h, sqrth, n1, n2, lz, ly);
    if (lep) {
        s->x = s->x - L * nx;
        s->y = s->y - L * ny;
    }
}

static void orc_simulate_impl(const orc_problem *pb, const orc_run *run, double x0, double y0, int geom,
                              const double *rect, uint32_t knot, int64_t path0, int64_t npaths,
                              double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps,
                              double *xt2, double *yt2, double *Y2, double *Z2, double *xi2, int64_t *steps2) {
    const double h = run->h, sqrth = sqrt(h);
    const int64_t cap = run->max_steps > 0 ? run->max_steps : INT64_MAX;
    const int rmask = geom ? 0 : (pb->refl & 15);
    if (rmask) {
        /* reflecting sides (rectangles): the path carries its nearest side and the local time of
           a projection back onto a reflecting side, and stops only on the absorbing ones */
#ifdef _OPENMP
        int nthr = run->threads > 0 ? run->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthr)
#endif
        for (int64_t n = 0; n < npaths; ++n) {
            uint32_t path = (uint32_t)(path0 + n);
            orc_state s;
            s.x = x0; s.y = y0; s.Y = 1.0; s.Z = 0.0; s.xi = 0.0;
            orc_coeffs(pb, &s);
            double ex, ey, rnx, rny, lt = 0.0;
            double rd = orc_rect_dist(rect, s.x, s.y, &ex, &ey, &rnx, &rny);
            int rk = orc_side_of_normal(rnx, rny);
            int64_t st = 0, stp = 0;
            for (int phase = 0; phase < (run->phase2 ? 2 : 1); ++phase) {
                const double beta = phase == 0 ? run->beta : 0.0;
                int64_t k = 0;
                int go;
                do {
                    double n1, n2, nx, ny;
                    orc_normal_pair(run, knot, path, st, &n1, &n2);
                    const int lep = orc_reflects(rk, rmask) && (pb->rho <= 0.0 || rd > -pb->rho);
                    const double E = lep ? orc_exp_draw(run, knot, path, st) : 0.0;
                    orc_refl_step(pb, &s, h, sqrth, n1, n2, lep, E, rnx, rny, rd, lt);
                    ++st; ++k;
                    double d = orc_rect_dist(rect, s.x, s.y, &ex, &ey, &nx, &ny);
                    int kind = orc_side_of_normal(nx, ny);
                    int rnow = orc_reflects(kind, rmask);
                    lt = 0.0;
                    if (rnow && d > 0.0) {
                        /* past a reflecting side: back onto it (Inside's reflect outcome,
                           Integrators/FKACIntegrator.hpp:189-207) */
                        lt = d;
                        s.x = ex; s.y = ey;
                        d = orc_rect_dist(rect, s.x, s.y, &ex, &ey, &nx, &ny);
                        kind = orc_side_of_normal(nx, ny);
                        rnow = orc_reflects(kind, rmask);
                    }
                    rd = d; rk = kind; rnx = nx; rny = ny;
                    orc_coeffs(pb, &s);
                    const double a = nx * s.sig[0] + ny * s.sig[2];
                    const double b = nx * s.sig[1] + ny * s.sig[3];
                    go = rnow || d < beta * sqrt(a * a + b * b) * sqrth;
                } while (go && k < cap);
                double *o[5] = {xt, yt, Y, Z, xi};
                int64_t *os = steps;
                if (phase == 1) { o[0] = xt2; o[1] = yt2; o[2] = Y2; o[3] = Z2; o[4] = xi2; os = steps2; }
                if (o[0]) o[0][n] = ex;
                if (o[1]) o[1][n] = ey;
                if (o[2]) o[2][n] = s.Y;
                if (o[3]) o[3][n] = s.Z;
                if (o[4]) o[4][n] = s.xi;
                if (os) os[n] = st - stp;
                stp = st;
            }
        }
        return;
    }
#ifdef _OPENMP
    int nthr = run->threads > 0 ? run->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthr)
#endif
    for (int64_t n = 0; n < npaths; ++n) {
        uint32_t path = (uint32_t)(path0 + n);
        orc_state s;
        s.x = x0; s.y = y0; s.Y = 1.0; s.Z = 0.0; s.xi = 0.0;
        orc_coeffs(pb, &s);
        int64_t st = 0;
        double ex = x0, ey = y0;
        int in;
        do {
            double n1, n2;
            orc_normal_pair(run, knot, path, st, &n1, &n2);
            orc_step(pb->step, &s, h, sqrth, n1, n2, 0.0, 0.0);
            ++st;
            in = orc_inside(pb, &s, geom, rect, run->beta, sqrth, &ex, &ey);
        } while (in && st < cap);
        if (xt) xt[n] = ex;
        if (yt) yt[n] = ey;
        if (Y) Y[n] = s.Y;
        if (Z) Z[n] = s.Z;
        if (xi) xi[n] = s.xi;
        if (steps) steps[n] = st;
        if (run->phase2) {
            /* sublinear continuation to the true exit (Solvers/EMFKACSolver.hpp:194-211) */
            int64_t st2 = 0;
            do {
                double n1, n2;
                orc_normal_pair(run, knot, path, st + st2, &n1, &n2);
                orc_step(pb->step, &s, h, sqrth, n1, n2, 0.0, 0.0);
                ++st2;
                in = orc_inside(pb, &s, geom, rect, 0.0, sqrth, &ex, &ey);
            } while (in && st2 < cap);
            if (xt2) xt2[n] = ex;
            if (yt2) yt2[n] = ey;
            if (Y2) Y2[n] = s.Y;
            if (Z2) Z2[n] = s.Z;
            if (xi2) xi2[n] = s.xi;
            if (steps2) steps2[n] = st2;
        }
    }
}

void orc_simulate(const orc_problem *pb, const orc_run *run, double x0, double y0, const double rect[4],
                  uint32_t knot, int64_t path0, int64_t npaths,
                  double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps,
                  double *xt2, double *yt2, double *Y2, double *Z2, double *xi2, int64_t *steps2) {
    orc_simulate_impl(pb, run, x0, y0, 0, rect, knot, path0, npaths, xt, yt, Y, Z, xi, steps, xt2, yt2, Y2, Z2,
                      xi2, steps2);
}

void orc_simulate_geom(const orc_problem *pb, const orc_run *run, double x0, double y0, int geom,
                       const double *dom, uint32_t knot, int64_t path0, int64_t npaths,
                       double *xt, double *yt, double *Y, double *Z, double *xi, int64_t *steps) {
    orc_simulate_impl(pb, run, x0, y0, geom, dom, knot, path0, npaths, xt, yt, Y, Z, xi, steps, NULL, NULL, NULL,
                      NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Reductions                                                                */
/* ------------------------------------------------------------------------ */
static inline int orc_same(double a, double b) { return fabs(a - b) < 1e-8; } /* Meshes/stencil.cpp:186-189 */

/* Exit side of a recorded exit point: 4 global boundary, 1 S, 0 N, 3 W, 2 E, -1 none.
   Global test of Reduce_Analytic (Solvers/EMFKACSolver.hpp:559) then the S,N,W,E priority of
   Stencil::G_update (Meshes/stencil.cpp:244-307). Side numbering N=0,S=1,E=2,W=3. */
static inline int orc_classify(const double rect[4], const double glob[4], double x, double y) {
    double ex, ey, nx, ny;
    if (orc_rect_dist(glob, x, y, &ex, &ey, &nx, &ny) >= -1e-8) return 4;
    if (orc_same(y, rect[1])) return 1;
    if (orc_same(y, rect[3])) return 0;
    if (orc_same(x, rect[0])) return 3;
    if (orc_same(x, rect[2])) return 2;
    return -1;
}

/* Exit side on a sector stencil dom of a disc of radius R (glob[2]): global if on the disc
   boundary, then inner arc, outer arc, ray a, ray b (the S, N, W, E priority of the rectangle). */
static inline int orc_classify_geom(int geom, const double *dom, const double glob[4], double x, double y) {
    if (!geom) return orc_classify(dom, glob, x, y);
    const double r = sqrt(x * x + y * y);
    if (r >= glob[2] - 1e-8) return 4;
    if (dom[0] > 0.0 && orc_same(r, dom[0])) return 1;
    if (orc_same(r, dom[1])) return 0;
    if (dom[0] < 0.0) return -1;
    if (fabs(dom[2] * y - dom[3] * x) < 1e-8) return 3;
    if (fabs(x * dom[5] - y * dom[4]) < 1e-8) return 2;
    return -1;
}

/* Gaussian RBF exp(-|p - X|^2 / c^2) with the reference's sqrt-then-square (BVPs/BVP.hpp:84-87) */
static inline double orc_rbf(double px, double py, double x, double y, double c) {
    double dx = px - x, dy = py - y;
    double r = sqrt(dx * dx + dy * dy);
    return orc_exp(-(r * r) / (c * c));
}

/* The same kernel as the GPU engine evaluates it: squared distance times a precomputed 1/c^2 */
static inline double orc_rbf_ic2(double px, double py, double x, double y, double inv_c2) {
    double dx = px - x, dy = py - y;
    return orc_exp_tab(-(dx * dx + dy * dy) * inv_c2);
}

/* Gaussian RBF values phi_j = exp(-|p_j - X|^2 / c^2) of the entries j of one straight,
   uniformly spaced stencil side (a rectangle side), chunk by chunk of ORC_RBF_CHUNK entries: the
   chunk's first entry j0 directly, the others by the two-multiply recurrence
   phi_{k+1} = phi_k r_k, r_{k+1} = r_k q, where with the side's step s (its first two entries)
   and d = p_j0 - X, q = exp(-2 |s|^2 / c^2) and r_0 = exp(-(2 d.s + |s|^2) / c^2) (above 1 while
   the entries approach X; its exponent is capped at 700, so an underflowed start stays 0).
   Replaces one exponential per entry (G_update's phi, Meshes/stencil.cpp:246-256) by two per chunk
   and one per side; the values equal the direct ones to about 1e-13 relative (values below
   about 1e-180 may come out as 0).  The engine's reduction evaluates exactly these operations. */
#define ORC_RBF_CHUNK 16
static void orc_rbf_side(const double *px, const double *py, int m, double x, double y, double ic2, double *phi) {
    double sx = 0.0, sy = 0.0, ss = 1.0;
    if (m > 1) {
        sx = px[1] - px[0];
        sy = py[1] - py[0];
        ss = sx * sx + sy * sy;
    }
    const double q = orc_exp_tab(-(2.0 * ss) * ic2);
    for (int j0 = 0; j0 < m; j0 += ORC_RBF_CHUNK) {
        const int j1 = j0 + ORC_RBF_CHUNK < m ? j0 + ORC_RBF_CHUNK : m;
        const double dx = px[j0] - x, dy = py[j0] - y;
        double v = orc_exp_tab(-(dx * dx + dy * dy) * ic2);
        double r = orc_exp(fmin(-(2.0 * (dx * sx + dy * sy) + ss) * ic2, 700.0));
        phi[j0] = v;
        for (int j = j0 + 1; j < j1; ++j) { v = v * r; phi[j] = v; r = r * q; }
    }
}

void orc_rbf_side_eval(const double *px, const double *py, int m, double x, double y, double ic2, double *phi) {
    orc_rbf_side(px, py, m, x, y, ic2, phi);
}

/* CSR row of a stencil: sorted unique columns; for each side entry its column slot. */
static int orc_row_pattern(const int msz[4], const int *idx, int *cols, int *slot) {
    int tot = msz[0] + msz[1] + msz[2] + msz[3];
    int n = 0;
    for (int e = 0; e < tot; ++e) {
        int v = idx[e], found = 0;
        for (int k = 0; k < n; ++k) if (cols[k] == v) { found = 1; break; }
        if (!found) cols[n++] = v;
    }
    /* insertion sort */
    for (int a = 1; a < n; ++a) {
        int v = cols[a], b = a - 1;
        while (b >= 0 && cols[b] > v) { cols[b + 1] = cols[b]; --b; }
        cols[b + 1] = v;
    }
    for (int e = 0; e < tot; ++e)
        for (int k = 0; k < n; ++k) if (cols[k] == idx[e]) { slot[e] = k; break; }
    return n;
}

int orc_reduce_faithful(const orc_problem *pb, int64_t npaths, const double *xt, const double *yt,
                        const double *Y, const double *Z, const double *xi,
                        const double rect[4], const double glob[4], const int msz[4],
                        const int *idx, const double *px, const double *py, const double cshape[4],
                        const double *ipsi, int quirk,
                        int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6]) {
    int off[5] = {0, msz[0], msz[0] + msz[1], msz[0] + msz[1] + msz[2], msz[0] + msz[1] + msz[2] + msz[3]};
    int64_t ioff[5];
    ioff[0] = 0;
    for (int s = 0; s < 4; ++s) ioff[s + 1] = ioff[s] + (int64_t)msz[s] * msz[s];
    int tot = off[4];
    double *Gs = (double *)calloc((size_t)tot + 1, sizeof(double));
    double *GGs = (double *)calloc((size_t)tot + 1, sizeof(double));
    double *phi = (double *)malloc(sizeof(double) * ((size_t)tot + 1));
    for (int k = 0; k < 6; ++k) cnt[k] = 0;
    double b = 0.0, bb = 0.0;
    for (int64_t n = 0; n < npaths; ++n) {
        int side = orc_classify(rect, glob, xt[n], yt[n]);
        if (side == 4) {
            double sig[4], bv[2], c, f, grad[2], g, u;
            orc_pde_eval(pb, xt[n], yt[n], sig, bv, &c, &f, grad, &g, &u);
            double sc = Z[n] + Y[n] * g + xi[n];
            b += sc; bb += sc * sc; cnt[4]++;
            continue;
        }
        if (side < 0) { cnt[5]++; }
        else {
            cnt[side]++;
            int m = msz[side];
            const double *A = ipsi + ioff[side];
            for (int j = 0; j < m; ++j)
                phi[j] = orc_rbf(px[off[side] + j], py[off[side] + j], xt[n], yt[n], cshape[side]);
            for (int i = 0; i < m; ++i) {
                double H = 0.0;
                for (int j = 0; j < m; ++j) H += A[(int64_t)j * m + i] * phi[j];
                Gs[off[side] + i] += -H * Y[n];
                GGs[off[side] + i] += (H * Y[n]) * (H * Y[n]);
            }
        }
        double sc = Z[n] + xi[n];
        b += sc; bb += sc * sc;
    }
    int *slot = (int *)malloc(sizeof(int) * ((size_t)tot + 1));
    int nnz = orc_row_pattern(msz, idx, cols, slot);
    double *gg = (double *)calloc((size_t)nnz + 1, sizeof(double));
    for (int k = 0; k < nnz; ++k) G[k] = 0.0;
    double norm = 1.0 / ((double)npaths * 1.0);
    /* G_return_withrep order: N, S, E, W (Meshes/stencil.cpp:373-396) */
    for (int e = 0; e < tot; ++e) {
        if (quirk) { G[slot[e]] = Gs[e] * norm; gg[slot[e]] = GGs[e] * norm; }
        else { G[slot[e]] += Gs[e] * norm; gg[slot[e]] += GGs[e] * norm; }
    }
    for (int k = 0; k < nnz; ++k) Gvar[k] = gg[k] - G[k] * G[k];
    *B = b / (double)npaths;
    *BB = bb / (double)npaths;
    free(Gs); free(GGs); free(phi); free(slot); free(gg);
    return nnz;
}

/* Sum over a row's columns of v[k] = G_b,k * w[k] in the engine's reduction order: column k to
   thread k mod 256 of a 256-thread work-group (each thread adding its columns in increasing k),
   a xor butterfly (offsets 32, 16, ..., 1) inside each 64-lane wave, then the four waves in
   order.  The order is the engine's own design, restated so that its sums compare bit for bit. */
static double orc_wg_sum(const double *G, const double *w, int nnz) {
    double part[256];
    for (int t = 0; t < 256; ++t) {
        part[t] = 0.0;
        for (int k = t; k < nnz; k += 256) part[t] = fma(G[k], w[k], part[t]);
    }
    double T = 0.0;
    for (int wv = 0; wv < 4; ++wv) {
        double v[64], nv[64];
        for (int i = 0; i < 64; ++i) v[i] = part[wv * 64 + i];
        for (int off = 32; off > 0; off >>= 1) {
            for (int i = 0; i < 64; ++i) nv[i] = v[i] + v[i ^ off];
            memcpy(v, nv, sizeof(v));
        }
        T = T + v[0];
    }
    return T;
}

int orc_reduce_batched(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt,
                       const double *yt, const double *Y, const double *Z, const double *xi,
                       const double rect[4], const double glob[4], const int msz[4],
                       const int *idx, const double *px, const double *py, const double cshape[4],
                       const double *ipsi, int quirk,
                       int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6]) {
    return orc_reduce_batched_eq(pb, npaths, S, P, xt, yt, Y, Z, xi, rect, glob, msz, idx, px, py, cshape, ipsi,
                                 quirk, NULL, cols, G, Gvar, B, BB, cnt, NULL, NULL);
}

/* As orc_reduce_batched; with w (per column of the row, the reference interface value of the
   column's knot) also the mean and batch-means per-path variance of the knot's equation estimator
   score - sum_k g_k w_k: per sub-batch R_b = (sum of scores) - sum_k G_b,k w_k. */
int orc_reduce_batched_eq(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt,
                          const double *yt, const double *Y, const double *Z, const double *xi,
                          const double rect[4], const double glob[4], const int msz[4],
                          const int *idx, const double *px, const double *py, const double cshape[4],
                          const double *ipsi, int quirk, const double *w,
                          int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6],
                          double *eq_mean, double *eq_var) {
    return orc_reduce_batched_geom(pb, npaths, S, P, xt, yt, Y, Z, xi, 0, rect, glob, msz, idx, px, py, cshape,
                                   ipsi, quirk, w, NULL, cols, G, Gvar, B, BB, cnt, eq_mean, eq_var);
}

/* As orc_reduce_batched_eq, with the exit sides classified on stopping domain dom of geometry
   geom (0 rectangle, 1 annular sector of a disc of radius glob[2]). */
int orc_reduce_batched_geom(const orc_problem *pb, int64_t npaths, int64_t S, int64_t P, const double *xt,
                            const double *yt, const double *Y, const double *Z, const double *xi, int geom,
                            const double *rect, const double glob[4], const int msz[4],
                            const int *idx, const double *px, const double *py, const double cshape[4],
                            const double *ipsi, int quirk, const double *w, const double *cent,
                            int *cols, double *G, double *Gvar, double *B, double *BB, int64_t cnt[6],
                            double *eq_mean, double *eq_var) {
    int off[5] = {0, msz[0], msz[0] + msz[1], msz[0] + msz[1] + msz[2], msz[0] + msz[1] + msz[2] + msz[3]};
    int64_t ioff[5];
    ioff[0] = 0;
    for (int s = 0; s < 4; ++s) ioff[s + 1] = ioff[s] + (int64_t)msz[s] * msz[s];
    int tot = off[4];
    int *slot = (int *)malloc(sizeof(int) * ((size_t)tot + 1));
    int nnz = orc_row_pattern(msz, idx, cols, slot);
    /* in quirk mode only the last side listing a column contributes to it */
    int *use = (int *)malloc(sizeof(int) * ((size_t)tot + 1));
    for (int e = 0; e < tot; ++e) {
        use[e] = 1;
        if (quirk) for (int e2 = e + 1; e2 < tot; ++e2) if (slot[e2] == slot[e]) { use[e] = 0; break; }
    }
    double ic2[4];
    for (int s = 0; s < 4; ++s) ic2[s] = 1.0 / (cshape[s] * cshape[s]);
    double *Phi = (double *)calloc((size_t)tot + 1, sizeof(double));
    double *phis = (double *)calloc((size_t)tot + 1, sizeof(double));
    double *Gb = (double *)calloc((size_t)nnz + 1, sizeof(double));
    double *isum = (double *)calloc((size_t)nnz + 1, sizeof(double));
    double *isq = (double *)calloc((size_t)nnz + 1, sizeof(double));
    double *sum = (double *)calloc((size_t)nnz + 1, sizeof(double));
    double *sq = (double *)calloc((size_t)nnz + 1, sizeof(double));
    for (int k = 0; k < 6; ++k) cnt[k] = 0;
    double b = 0.0, bb = 0.0, rs = 0.0, rq = 0.0;
    int64_t nb = 0;
    /* two-level order: sub-batches of S paths summed into items of P paths, items summed into the knot */
    for (int64_t w0 = 0; w0 < npaths; w0 += P) {
        int64_t w1 = w0 + P < npaths ? w0 + P : npaths;
        double ib = 0.0, ibb = 0.0, irs = 0.0, irq = 0.0;
        memset(isum, 0, sizeof(double) * (size_t)nnz);
        memset(isq, 0, sizeof(double) * (size_t)nnz);
        for (int64_t p0 = w0; p0 < w1; p0 += S) {
            int64_t p1 = p0 + S < w1 ? p0 + S : w1;
            memset(Phi, 0, sizeof(double) * (size_t)tot);
            double sb = 0.0, sbb = 0.0;
            for (int64_t n = p0; n < p1; ++n) {
                int side = orc_classify_geom(geom, rect, glob, xt[n], yt[n]);
                double sc;
                if (side == 4) {
                    double sig[4], bv[2], c, f, grad[2], g, u;
                    orc_pde_eval(pb, xt[n], yt[n], sig, bv, &c, &f, grad, &g, &u);
                    sc = Z[n] + Y[n] * g + xi[n];
                    cnt[4]++;
                } else {
                    if (side < 0) cnt[5]++;
                    else if (geom == 0) {
                        /* rectangle sides: straight and uniformly spaced, one Gaussian width (the
                           side's, or its first entry's when widths are given per entry) */
                        cnt[side]++;
                        const double ics = cent ? 1.0 / (cent[off[side]] * cent[off[side]]) : ic2[side];
                        orc_rbf_side(px + off[side], py + off[side], msz[side], xt[n], yt[n], ics, phis);
                        for (int j = 0; j < msz[side]; ++j)
                            Phi[off[side] + j] = fma(Y[n], phis[j], Phi[off[side] + j]);
                    } else {
                        cnt[side]++;
                        for (int j = 0; j < msz[side]; ++j)
                            Phi[off[side] + j] = fma(Y[n],
                                orc_rbf_ic2(px[off[side] + j], py[off[side] + j], xt[n], yt[n],
                                            cent ? 1.0 / (cent[off[side] + j] * cent[off[side] + j]) : ic2[side]),
                                Phi[off[side] + j]);
                    }
                    sc = Z[n] + xi[n];
                }
                sb = sb + sc;
                sbb = fma(sc, sc, sbb);
            }
            memset(Gb, 0, sizeof(double) * (size_t)nnz);
            for (int s = 0; s < 4; ++s) {
                int m = msz[s];
                const double *A = ipsi + ioff[s];
                for (int i = 0; i < m; ++i) {
                    int e = off[s] + i;
                    if (!use[e]) continue;
                    double H = 0.0;
                    for (int j = 0; j < m; ++j) H = fma(A[(int64_t)j * m + i], Phi[off[s] + j], H);
                    Gb[slot[e]] = Gb[slot[e]] + (-H);
                }
            }
            double nbeta = (double)(p1 - p0);
            for (int k = 0; k < nnz; ++k) { isum[k] = isum[k] + Gb[k]; isq[k] = isq[k] + (Gb[k] * Gb[k]) / nbeta; }
            ib = ib + sb;
            ibb = ibb + sbb;
            if (w) {
                const double R = sb - orc_wg_sum(Gb, w, nnz);
                irs = irs + R;
                irq = irq + (R * R) / nbeta;
            }
            nb++;
        }
        for (int k = 0; k < nnz; ++k) { sum[k] = sum[k] + isum[k]; sq[k] = sq[k] + isq[k]; }
        b = b + ib;
        bb = bb + ibb;
        rs = rs + irs;
        rq = rq + irq;
    }
    double N = (double)npaths;
    for (int k = 0; k < nnz; ++k) {
        G[k] = sum[k] / N;
        Gvar[k] = nb > 1 ? (sq[k] - N * G[k] * G[k]) / (double)(nb - 1) : 0.0;
    }
    *B = b / N;
    *BB = bb / N;
    if (w && eq_mean && eq_var) {
        *eq_mean = rs / N;
        *eq_var = nb > 1 ? (rq - N * *eq_mean * *eq_mean) / (double)(nb - 1) : 0.0;
    }
    free(slot); free(use); free(Phi); free(phis); free(Gb); free(isum); free(isq); free(sum); free(sq);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Psi inverse by full-pivot LU (Meshes/stencil.cpp:115-184)                 */
/* ------------------------------------------------------------------------ */
int orc_stencil_ipsi(int m, const double *px, const double *py, double c, double *ipsi, double *cond) {
    double *A = (double *)malloc(sizeof(double) * (size_t)m * m);
    double *LU = (double *)malloc(sizeof(double) * (size_t)m * m);
    int *pr = (int *)malloc(sizeof(int) * (size_t)m), *pc = (int *)malloc(sizeof(int) * (size_t)m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[(int64_t)i * m + j] = orc_rbf(px[i], py[i], px[j], py[j], c);
    memcpy(LU, A, sizeof(double) * (size_t)m * m);
    for (int i = 0; i < m; ++i) { pr[i] = i; pc[i] = i; }
    double maxpiv = 0.0;
    int rank = m;
    for (int k = 0; k < m; ++k) {
        int bi = k, bj = k; double best = -1.0;
        for (int i = k; i < m; ++i)
            for (int j = k; j < m; ++j) { double v = fabs(LU[(int64_t)i * m + j]); if (v > best) { best = v; bi = i; bj = j; } }
        if (k == 0) maxpiv = best;
        if (best <= maxpiv * 2.220446049250313e-16 * m) { rank = k; break; }
        if (bi != k) { for (int j = 0; j < m; ++j) { double t = LU[(int64_t)k * m + j]; LU[(int64_t)k * m + j] = LU[(int64_t)bi * m + j]; LU[(int64_t)bi * m + j] = t; } int t = pr[k]; pr[k] = pr[bi]; pr[bi] = t; }
        if (bj != k) { for (int i = 0; i < m; ++i) { double t = LU[(int64_t)i * m + k]; LU[(int64_t)i * m + k] = LU[(int64_t)i * m + bj]; LU[(int64_t)i * m + bj] = t; } int t = pc[k]; pc[k] = pc[bj]; pc[bj] = t; }
        double piv = LU[(int64_t)k * m + k];
        for (int i = k + 1; i < m; ++i) {
            double l = LU[(int64_t)i * m + k] / piv;
            LU[(int64_t)i * m + k] = l;
            for (int j = k + 1; j < m; ++j) LU[(int64_t)i * m + j] -= l * LU[(int64_t)k * m + j];
        }
    }
    int status = 0;
    if (rank < m) status = 1;
    else {
        /* P A Q = L U  =>  A^-1 = Q U^-1 L^-1 P */
        double *col = (double *)malloc(sizeof(double) * (size_t)m);
        for (int e = 0; e < m; ++e) {
            /* solve A z = e_e : L U (Q^T z) = P e_e */
            for (int i = 0; i < m; ++i) col[i] = (pr[i] == e) ? 1.0 : 0.0;
            for (int i = 0; i < m; ++i) { double s = col[i]; for (int j = 0; j < i; ++j) s -= LU[(int64_t)i * m + j] * col[j]; col[i] = s; }
            for (int i = m - 1; i >= 0; --i) { double s = col[i]; for (int j = i + 1; j < m; ++j) s -= LU[(int64_t)i * m + j] * col[j]; col[i] = s / LU[(int64_t)i * m + i]; }
            for (int i = 0; i < m; ++i) ipsi[(int64_t)pc[i] * m + e] = col[i];
        }
        free(col);
    }
    double na = 0.0, ni = 0.0;
    for (int64_t k = 0; k < (int64_t)m * m; ++k) { na += A[k] * A[k]; if (!status) ni += ipsi[k] * ipsi[k]; }
    *cond = status ? INFINITY : sqrt(na) * sqrt(ni);
    free(A); free(LU); free(pr); free(pc);
    return status;
}

/* ------------------------------------------------------------------------ */
/* BiCGSTAB (van der Vorst 1992) on (I + G): the caller passes the rows of G */
/* without the identity (Compute_Solution adds it, PDDSparseGM.cpp:3187-3189). */
/* Zero initial guess, stop when |r| <= tol |b| or maxit iterations          */
/* (PDDAlgorithms/PDDSparseGM.cpp:3228-3232).                                */
/* ------------------------------------------------------------------------ */
static void orc_spmv(int n, const int64_t *rp, const int *ci, const double *v, const double *x, double *y) {
    for (int i = 0; i < n; ++i) { double s = x[i]; for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]]; y[i] = s; }
}
static double orc_dot(int n, const double *a, const double *b) { double s = 0.0; for (int i = 0; i < n; ++i) s += a[i] * b[i]; return s; }

int orc_bicgstab(int n, const int64_t *rp, const int *ci, const double *v, const double *b,
                 double *x, double tol, int maxit, double *relres) {
    double *r = calloc((size_t)n, 8), *r0 = calloc((size_t)n, 8), *p = calloc((size_t)n, 8), *vv = calloc((size_t)n, 8);
    double *s = calloc((size_t)n, 8), *t = calloc((size_t)n, 8);
    for (int i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; r0[i] = b[i]; }
    double bn2 = orc_dot(n, b, b);
    if (bn2 == 0.0) { *relres = 0.0; free(r); free(r0); free(p); free(vv); free(s); free(t); return 0; }
    double rho = 1.0, alpha = 1.0, w = 1.0, r0n2 = bn2;
    int it = 0;
    double tol2 = tol * tol * bn2;
    while (orc_dot(n, r, r) > tol2 && it < maxit) {
        double rho_old = rho;
        rho = orc_dot(n, r0, r);
        if (fabs(rho) < 1e-30 * r0n2) { /* breakdown: restart with the current residual */
            orc_spmv(n, rp, ci, v, x, t);
            for (int i = 0; i < n; ++i) { r[i] = b[i] - t[i]; r0[i] = r[i]; }
            rho = r0n2 = orc_dot(n, r, r);
        }
        double beta = (rho / rho_old) * (alpha / w);
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - w * vv[i]);
        orc_spmv(n, rp, ci, v, p, vv);
        alpha = rho / orc_dot(n, r0, vv);
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * vv[i];
        orc_spmv(n, rp, ci, v, s, t);
        double tt = orc_dot(n, t, t);
        w = tt > 0.0 ? orc_dot(n, t, s) / tt : 0.0;
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i] + w * s[i]; r[i] = s[i] - w * t[i]; }
        ++it;
    }
    *relres = sqrt(orc_dot(n, r, r) / bn2);
    free(r); free(r0); free(p); free(vv); free(s); free(t);
    return it;
}

/* ------------------------------------------------------------------------ */
/* LUT prefilter: interpolating cubic B-spline coefficients                  */
/* ------------------------------------------------------------------------ */
/* 1-D: given s[0..n-1] (stride ss) find c[-1..n] (n+2 values, stride cs) with
   (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k], and end conditions matching the
   spline's second difference to the data's, extrapolated to the end node. */
static void orc_prefilter_1d(int n, const double *s, int ss, double *c, int cs, double *work) {
    double rL = 2.0 * (s[0] - 2.0 * s[ss] + s[2 * ss]) - (s[ss] - 2.0 * s[2 * ss] + s[3 * ss]);
    double rR = 2.0 * (s[(n - 1) * ss] - 2.0 * s[(n - 2) * ss] + s[(n - 3) * ss])
                - (s[(n - 2) * ss] - 2.0 * s[(n - 3) * ss] + s[(n - 4) * ss]);
    double c0 = s[0] - rL / 6.0, cn1 = s[(n - 1) * ss] - rR / 6.0;
    /* unknowns c[1..n-2]: c[k-1] + 4 c[k] + c[k+1] = 6 s[k] */
    int m = n - 2;
    double *cp = work, *dp = work + m;
    for (int k = 0; k < m; ++k) {
        double rhs = 6.0 * s[(k + 1) * ss];
        if (k == 0) rhs -= c0;
        if (k == m - 1) rhs -= cn1;
        double denom = 4.0 - (k > 0 ? cp[k - 1] : 0.0);
        cp[k] = 1.0 / denom;
        dp[k] = (rhs - (k > 0 ? dp[k - 1] : 0.0)) / denom;
    }
    double *cc = (double *)malloc(sizeof(double) * (size_t)n);
    cc[0] = c0; cc[n - 1] = cn1;
    for (int k = m - 1; k >= 0; --k) cc[k + 1] = dp[k] - cp[k] * (k + 1 < m ? cc[k + 2] : 0.0);
    c[0] = 2.0 * cc[0] - cc[1] + rL;
    for (int k = 0; k < n; ++k) c[(k + 1) * cs] = cc[k];
    c[(n + 1) * cs] = 2.0 * cc[n - 1] - cc[n - 2] + rR;
    free(cc);
}

void orc_lut_prefilter(int nx, int ny, int nf, const double *samples, float *coef) {
    int NX = nx + 2, NY = ny + 2;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)NX * ny);
    double *out = (double *)malloc(sizeof(double) * (size_t)NX * NY);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)(nx > ny ? nx : ny));
    for (int f = 0; f < nf; ++f) {
        double *row = (double *)malloc(sizeof(double) * (size_t)nx);
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) row[i] = samples[((int64_t)j * nx + i) * nf + f];
            orc_prefilter_1d(nx, row, 1, tmp + (int64_t)j * NX, 1, work);
        }
        free(row);
        double *colv = (double *)malloc(sizeof(double) * (size_t)ny);
        double *colc = (double *)malloc(sizeof(double) * (size_t)NY);
        for (int i = 0; i < NX; ++i) {
            for (int j = 0; j < ny; ++j) colv[j] = tmp[(int64_t)j * NX + i];
            orc_prefilter_1d(ny, colv, 1, colc, 1, work);
            for (int j = 0; j < NY; ++j) out[(int64_t)j * NX + i] = colc[j];
        }
        free(colv); free(colc);
        for (int64_t k = 0; k < (int64_t)NX * NY; ++k) coef[k * nf + f] = (float)out[k];
    }
    free(tmp); free(out); free(work);
}
