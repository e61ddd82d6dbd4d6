// pdds_pde.h -- device coefficient functors: the compile-time replacement of the
// reference's struct bvp function pointers (BVPs/BVP.hpp:89-137) and of its
// compile-time device twin (BVPs/BVP.cuh:44-77, BVPs/initBVPdev.cuh:6-16).
//
// Reference convention: 1/2 s s^T : D2u + b . Du + c u + f = 0 with c <= 0,
// Dirichlet g = u on the boundary.  Each functor returns sigma, b, c, f and
// grad u (for the control variate xi) at a point, and g there.
#pragma once
#include "pdds_math.h"

namespace pdds {

enum PdeId { PDE_MONEGROS = 0, PDE_SINSIN = 1, PDE_VARCOEF = 2, PDE_VARCOEF_LUT = 3, PDE_HELMHOLTZ = 4, PDE_TABLE = 5 };
// Step families, the values of pdds_run.step.  Kernels are compiled for STEP_CV alone (the
// default step, no family test in the loop) and for STEP_ANY (the family read at run time).
enum StepFamily { STEP_ANY = -1, STEP_CV = 0, STEP_PLAIN = 1, STEP_VR = 2, STEP_VRCV = 3 };

template <class R> struct Coef {
    R s00, s01, s10, s11;  // sigma, row-major
    R b0, b1;
    R c, f;
    R g0, g1;              // grad u
    R m0, m1, F0, F1;      // STEP_VRCV: the drift change mu and the control variate's F
};

struct PdeParams {
    double k2;             // HELMHOLTZ (launches without per-knot values)
    LutView lut;           // VARCOEF_LUT: sigma11, sigma22, b1, b2, c; TABLE: those, f, du/dx, du/dy
    LutView ulut;          // TABLE: u (the Dirichlet data g, and u of the VR step)
    LutView vlut;          // STEP_VRCV: mu1, mu2, F1, F2
    int32_t step;          // StepFamily: PLAIN takes grad u as 0, VR hands out grad u / u
    int32_t refl;          // bit s: global side s (0 N, 1 S, 2 E, 3 W) reflects (Neumann / Robin)
    double robin;          // varphi on reflecting sides (<= 0); psi = grad u . N - varphi u
    double rho;            // Lepingle step within this distance of a reflecting side
};

// Monegros Poisson (BVPs/EDPs/Monegros_Poisson_CPP.hpp:4-81): u, grad u, Laplacian
template <class R>
PDDS_HD void monegros(R x, R y, R &u, R &gx, R &gy, R &lap) {
    const R K1 = (R)1.0, K2 = (R)0.01, K3 = (R)0.02, K4 = (R)0.12, K5 = (R)0.05, K6 = (R)0.05, K7 = (R)-0.12;
    const R SF = (R)0.333;
    R a1 = RM<R>::sqrt(K1 + K2 * x * x + K3 * y * y);
    R s1, c1, s2, c2, sa, ca;
    RM<R>::sincos(K4 * x + K5 * y, s1, c1);
    RM<R>::sincos(K6 * x + K7 * y, s2, c2);
    R T = RM<R>::tanh(s1 + s2);
    RM<R>::sincos(a1, sa, ca);
    R T21 = T * T - (R)1.0;
    R px = K4 * c1 + K6 * c2, py = K5 * c1 + K7 * c2;
    R a3 = a1 * a1 * a1;
    u = (R)3.0 + SF * (sa + T);
    gx = SF * (RM<R>::div(ca * K2 * x, a1) - px * T21);
    gy = SF * (RM<R>::div(ca * K1 * K3 * y, a1) - py * T21);
    R dxx = SF * (((RM<R>::div(K2, a1)) - RM<R>::div(K2 * K2 * x * x, a3)) * ca - (RM<R>::div(x * x * K2 * K2, a1 * a1)) * sa
                  + (R)2.0 * px * px * T21 * T + (K4 * K4 * s1 + K6 * K6 * s2) * T21);
    R dyy = SF * (((RM<R>::div(K3, a1)) - RM<R>::div(K3 * K3 * y * y, a3)) * ca - (RM<R>::div(y * y * K3 * K3, a1 * a1)) * sa
                  + (R)2.0 * py * py * T21 * T + (K5 * K5 * s1 + K7 * K7 * s2) * T21);
    lap = dxx + dyy;
}

template <int PDE, class R> PDDS_HD R pde_g(R x, R y, const PdeParams &P);

// TILED: the coefficient LUT's fields 0..3 are read from the LDS tile T where it covers X.
// k2: the Helmholtz parameter of this path's problem (per knot, so one launch can carry
// requests with different k^2).
template <int PDE, class R, bool TILED, int FAM = STEP_CV>
PDDS_HD void pde_coef_t(R x, R y, const PdeParams &P, const LutTile &T, R k2, Coef<R> &co) {
    (void)T;
    if constexpr (PDE == PDE_MONEGROS) {
        R u, gx, gy, lap;
        monegros<R>(x, y, u, gx, gy, lap);
        co.s00 = (R)1.41421356237; co.s01 = 0; co.s10 = 0; co.s11 = (R)1.41421356237;
        co.b0 = 0; co.b1 = 0; co.c = (R)-0.0; co.f = -lap;
        co.g0 = gx; co.g1 = gy;
    } else if constexpr (PDE == PDE_SINSIN || PDE == PDE_HELMHOLTZ) {
        R sx, cx, sy, cy;
        RM<R>::sincospi(x, sx, cx);
        RM<R>::sincospi(y, sy, cy);
        co.s00 = (R)1.4142135623730951; co.s01 = 0; co.s10 = 0; co.s11 = (R)1.4142135623730951;
        co.b0 = 0; co.b1 = 0;
        co.g0 = (R)3.141592653589793 * cx * sy + (R)2.0 * x;
        co.g1 = (R)3.141592653589793 * sx * cy + (R)2.0 * y;
        R flap = (R)19.739208802178716 * sx * sy - (R)4.0;
        if constexpr (PDE == PDE_SINSIN) {
            co.c = 0; co.f = flap;
        } else {
            R uu = sx * sy + x * x + y * y;
            co.c = -k2; co.f = flap + k2 * uu;
        }
    } else if constexpr (PDE == PDE_TABLE) {
        // every coefficient tabulated (the reference's texMode 4 serves f, c and grad u from
        // textures, Solvers/MCinCUDA.cu:453-478; here sigma and b as well), so a problem runs
        // from its samples alone
        R v[8];
        lut_eval<R, 8>(P.lut, x, y, v);
        co.s00 = v[0]; co.s01 = 0; co.s10 = 0; co.s11 = v[1];
        co.b0 = v[2]; co.b1 = v[3]; co.c = v[4]; co.f = v[5];
        co.g0 = v[6]; co.g1 = v[7];
    } else {  // VARCOEF, VARCOEF_LUT
        R sxn, cxn, syn, cyn;
        RM<R>::sincos(x, sxn, cxn);
        RM<R>::sincos(y, syn, cyn);
        R ex = RM<R>::exp(x);
        R s1 = (R)1.0 + (R)0.3 * sxn, s2 = (R)1.0 + (R)0.3 * cyn;
        R uu = ex * cyn, uy = -ex * syn;
        R a1 = s1 * s1, a2 = s2 * s2;
        R a1p = (R)2.0 * s1 * ((R)0.3 * cxn), a2p = (R)2.0 * s2 * ((R)-0.3 * syn);
        R fcfg = (R)0.5 * (a1p * uu + a1 * uu + a2p * uy - a2 * uu)
                 + (R)0.2 * y * uu + ((R)-0.2 * x) * uy - ((R)1.0 + x * x) * uu;
        co.f = -fcfg;
        co.g0 = uu; co.g1 = uy;
        if constexpr (PDE == PDE_VARCOEF) {
            co.s00 = s1; co.s01 = 0; co.s10 = 0; co.s11 = s2;
            co.b0 = (R)0.2 * y + (R)0.3 * s1 * cxn;
            co.b1 = (R)-0.2 * x - (R)0.3 * s2 * syn;
            co.c = -((R)1.0 + x * x);
        } else {
            R v[5];
            lut_eval_t<R, 5, TILED>(P.lut, T, x, y, v);
            co.s00 = v[0]; co.s01 = 0; co.s10 = 0; co.s11 = v[1];
            co.b0 = v[2]; co.b1 = v[3]; co.c = v[4];
        }
    }
    if constexpr (FAM != STEP_CV) {
        if (P.step == STEP_PLAIN) {
            // no control variate (Integrators/FKACIntegrator.hpp:20-29)
            co.g0 = 0;
            co.g1 = 0;
        } else if (P.step == STEP_VR) {
            // grad u / u, so that the step's -sigma^T g is mu (Integrators/FKACIntegrator.hpp:31-44)
            const R uu = pde_g<PDE, R>(x, y, P);     // u (every shipped problem has g = u)
            co.g0 = co.g0 / uu;
            co.g1 = co.g1 / uu;
        } else if (P.step == STEP_VRCV) {
            // mu and F from their table (Integrators/FKACIntegrator.hpp:100-112)
            R w[4];
            lut_eval<R, 4>(P.vlut, x, y, w);
            co.m0 = w[0]; co.m1 = w[1]; co.F0 = w[2]; co.F1 = w[3];
        }
    }
}

template <int PDE, class R, int FAM = STEP_CV>
PDDS_HD void pde_coef(R x, R y, const PdeParams &P, Coef<R> &co) {
    const LutTile none{nullptr, 0, 0, 0, 0};
    pde_coef_t<PDE, R, false, FAM>(x, y, P, none, (R)P.k2, co);
}

// Boundary data of a reflecting side with outward normal (nx, ny), the reference's psi and varphi
// (BVPs/BVP.hpp:125-130; Z += psi Y L and Y += varphi Y L per local-time increment L,
// Integrators/FKACIntegrator.hpp:156-158): the Robin condition du/dn = varphi u + psi of the
// manufactured solution, varphi the problem's constant (0: Neumann), so psi = grad u . n - varphi u.
template <int PDE, class R> PDDS_HD R pde_g(R x, R y, const PdeParams &P);
template <int PDE, class R>
PDDS_HD R pde_psi(R x, R y, R nx, R ny, R varphi, const PdeParams &P) {
    R gx, gy;
    if constexpr (PDE == PDE_MONEGROS) {
        R u, lap;
        monegros<R>(x, y, u, gx, gy, lap);
    } else if constexpr (PDE == PDE_SINSIN || PDE == PDE_HELMHOLTZ) {
        R sx, cx, sy, cy;
        RM<R>::sincospi(x, sx, cx);
        RM<R>::sincospi(y, sy, cy);
        gx = (R)3.141592653589793 * cx * sy + (R)2.0 * x;
        gy = (R)3.141592653589793 * sx * cy + (R)2.0 * y;
    } else if constexpr (PDE == PDE_TABLE) {
        R v[8];
        lut_eval<R, 8>(P.lut, x, y, v);
        gx = v[6];
        gy = v[7];
    } else {
        R s, c;
        RM<R>::sincos(y, s, c);
        const R ex = RM<R>::exp(x);
        gx = ex * c;
        gy = -ex * s;
    }
    return (gx * nx + gy * ny) - varphi * pde_g<PDE, R>(x, y, P);
}

// Dirichlet data g = u (the compiled-in problems are manufactured solutions; TABLE reads u from
// its one-field table)
template <int PDE, class R>
PDDS_HD R pde_g(R x, R y, const PdeParams &P) {
    (void)P;
    if constexpr (PDE == PDE_TABLE) {
        R u[1];
        lut_eval<R, 1>(P.ulut, x, y, u);
        return u[0];
    } else if constexpr (PDE == PDE_MONEGROS) {
        R u, gx, gy, lap;
        monegros<R>(x, y, u, gx, gy, lap);
        return u;
    } else if constexpr (PDE == PDE_SINSIN || PDE == PDE_HELMHOLTZ) {
        R sx, cx, sy, cy;
        RM<R>::sincospi(x, sx, cx);
        RM<R>::sincospi(y, sy, cy);
        return sx * sy + x * x + y * y;
    } else {
        R s, c;
        RM<R>::sincos(y, s, c);
        return RM<R>::exp(x) * c;
    }
}

}  // namespace pdds
