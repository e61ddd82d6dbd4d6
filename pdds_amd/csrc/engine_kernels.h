// engine_kernels.h -- HIP kernels of the path engine (templated on path precision R
// and on the PDE functor).  Included by engine_f64.hip (compiled with
// -ffp-contract=off, bit-exact with the FP64 oracle) and engine_f32.hip.
//
// Two kernels per assembly batch:
//  * path_kernel: one path per lane.  In persistent form every lane whose path
//    has stopped takes the next path of the batch from a device-wide counter
//    (one atomic per wave per refill), so lanes idle only at the very end of the
//    batch -- the on-device form of the reference's relaunch-and-refill loop
//    (Solvers/SolveLoop.cuh:59-74, solveCUDA.cuh:118-168).  Each path is stepped
//    by the control-variate Euler-Maruyama scheme (Integrators/FKACIntegrator.hpp:46-57),
//    stopped by the Gobet-shifted test (Integrators/FKACIntegrator.hpp:168-186),
//    and its exit record (projection, Y, score) is written at its global path
//    index; steps are summed per knot.  The plain form launches one lane per path,
//    no refill.
//  * reduce_kernel: one work-group per work item of P paths of one knot.  Per
//    sub-batch of S paths it forms the per-side RBF moments Phi_j = sum Y phi_j(X)
//    (m exps per exit, where Meshes/stencil.cpp:244-307 spends m^2), then
//    G_beta = -Psi^-T Phi per CSR column, and accumulates sums and sums of squares
//    (batch-means variance) into per-item partials.  finalize_kernel adds the
//    items of each knot in item order.  Every sum has a fixed order: results do
//    not depend on scheduling.
#pragma once
#include "pdds_pde.h"
#include "engine_types.h"

// Waves per SIMD the path kernels are register-budgeted for (a wave gets up to 512 / waves
// VGPRs).  FP64: the LUT kernels keep a step's LUT reads in flight with up to 256 VGPRs (2
// waves), as does the Monegros problem (three transcendental chains per step); the other
// variable-coefficient problem fits 3 waves and the Poisson/Helmholtz problems 4.  FP32: 3
// waves for the LUT problem, 2 for the tabulated problem (eight fields per lookup), 4 for the
// others.
template <class R, int PDE> constexpr int path_waves() {
    if constexpr (sizeof(R) == 4) return PDE == pdds::PDE_TABLE ? 2 : PDE == pdds::PDE_VARCOEF_LUT ? 3 : 4;
    else if constexpr (PDE == pdds::PDE_VARCOEF_LUT || PDE == pdds::PDE_MONEGROS || PDE == pdds::PDE_TABLE) return 2;
    else if constexpr (PDE == pdds::PDE_VARCOEF) return 3;
    else return 4;
}

namespace pdds {

// Exit classification of FP64 paths: global-boundary test of Reduce_Analytic
// (Solvers/EMFKACSolver.hpp:559) then the S,N,W,E coordinate test of
// Stencil::G_update (Meshes/stencil.cpp:246-301, tolerance 1e-8).
PDDS_HD int classify_coord(const double rect[4], const double glob[4], double x, double y) {
    double ex, ey, nx, ny;
    int kind;
    if (rect_dist<double>(glob, x, y, ex, ey, nx, ny, kind) >= -1e-8) return 4;
    if (__builtin_fabs(y - rect[1]) < 1e-8) return 1;
    if (__builtin_fabs(y - rect[3]) < 1e-8) return 0;
    if (__builtin_fabs(x - rect[0]) < 1e-8) return 3;
    if (__builtin_fabs(x - rect[2]) < 1e-8) return 2;
    return -1;
}

// classify_coord with the global rectangle given by its centre and half-sides gc = {cx, cy, hx, hy}
// (rect_dist's own cx = 0.5 (x0 + x1), hx = x1 - cx): rect_dist's value is max(|x - cx| - hx,
// |y - cy| - hy) inside the rectangle (its tie rule picks between equal values) and positive
// outside, where that maximum is positive too, so for finite coordinates the global test is the
// same comparison without the projection.
PDDS_HD int classify_coord_c(const double rect[4], const double gc[4], double x, double y) {
    const double ax = __builtin_fabs(x - gc[0]) - gc[2], ay = __builtin_fabs(y - gc[1]) - gc[3];
    if (__builtin_fmax(ax, ay) >= -1e-8) return 4;
    if (__builtin_fabs(y - rect[1]) < 1e-8) return 1;
    if (__builtin_fabs(y - rect[3]) < 1e-8) return 0;
    if (__builtin_fabs(x - rect[0]) < 1e-8) return 3;
    if (__builtin_fabs(x - rect[2]) < 1e-8) return 2;
    return -1;
}

// Exit classification of FP64 paths on an annular-sector stencil g (sector_dist) of a disc of
// radius R centred at the origin: the global test (on the disc boundary) first, then inner arc,
// outer arc, ray a, ray b -- the order of the rectangle's S, N, W, E test.
PDDS_HD int classify_sector(const double g[6], double R, double x, double y) {
    const double r = __builtin_sqrt(x * x + y * y);
    if (r >= R - 1e-8) return 4;
    if (g[0] > 0.0 && __builtin_fabs(r - g[0]) < 1e-8) return 1;
    if (__builtin_fabs(r - g[1]) < 1e-8) return 0;
    if (g[0] < 0.0) return -1;   // the whole disc r <= r1 has no other side
    if (__builtin_fabs(g[2] * y - g[3] * x) < 1e-8) return 3;
    if (__builtin_fabs(x * g[5] - y * g[4]) < 1e-8) return 2;
    return -1;
}

// Exit classification of FP32 paths from the integer projection kind: a side of
// the stencil rectangle that lies on the global boundary (mask bit s) gives a
// global exit; a corner is global if either adjacent side is, else it takes the
// horizontal side (the S/N-first priority of the coordinate test).
PDDS_HD int classify_kind(int kind, int gmask) {
    if (kind < 4) return ((gmask >> kind) & 1) ? 4 : kind;
    int v = (kind == 4 || kind == 5) ? 0 : 1;       // N for NE/NW, S for SE/SW
    int hs = (kind == 4 || kind == 6) ? 2 : 3;      // E for NE/SE, W for NW/SW
    if (((gmask >> v) & 1) || ((gmask >> hs) & 1)) return 4;
    return v;
}

// Gaussian RBF exp(-|p - X|^2 / c^2) (BVPs/BVP.hpp:84-87) with 1/c^2 precomputed; the table
// driven exponential with its 2^(j/64) table T
PDDS_HD double rbf_tab(double px, double py, double x, double y, double inv_c2, const double *T) {
    double dx = px - x, dy = py - y;
    return dexp_tab_nonpos(-(dx * dx + dy * dy) * inv_c2, T);
}

// The same Gaussian at exit (x, y) for the kRbfChunk entries from j0 of a straight, uniformly
// spaced stencil side (px, py; a rectangle side) with step (sx, sy), ss = |s|^2 and
// q = exp(-2 ss / c^2): entry j0 directly, the others by the two-multiply recurrence
// phi_{k+1} = phi_k r_k, r_{k+1} = r_k q with d = p_j0 - X and r_0 = exp(-(2 d.s + ss) / c^2)
// (above 1 while the entries approach X; its exponent capped at 700, so an underflowed start
// stays 0).  Two exponentials per chunk instead of one per entry; the values equal the direct ones
// to about 1e-13 relative (those below about 1e-180 may come out as 0).  Each value phi_k is added
// as acc[k] = fma(Y, phi_k, acc[k]); the sums for k past the side's end are formed and not used.
PDDS_HD void rbf_side_chunk(const double *px, const double *py, int j0, double sx, double sy, double ss, double q,
                            double x, double y, double inv_c2, const double *T, double Y, double acc[kRbfChunk]) {
    const double dx = px[j0] - x, dy = py[j0] - y;
    double p = dexp_tab_nonpos(-(dx * dx + dy * dy) * inv_c2, T);
    double r = dexp_fast(__builtin_fmin(-(2.0 * (dx * sx + dy * sy) + ss) * inv_c2, 700.0));
    acc[0] = __builtin_fma(Y, p, acc[0]);
#pragma unroll
    for (int k = 1; k < kRbfChunk; ++k) {
        p = p * r;
        acc[k] = __builtin_fma(Y, p, acc[k]);
        r = r * q;
    }
}

// ----------------------------------------------------------------------------
// Per-lane path state
// ----------------------------------------------------------------------------
template <class R> struct PathState {
    R x, y, Y, Z, xi;
    R zc, xc;          // Kahan compensations (FP32 only)
    R pn1, pn2;        // pending normals of the odd step (FP32 only)
    uint32_t st;       // steps taken (below 2^31: the phase-1 cap is below 2^30)
    uint32_t path;     // path index within its knot
};

// One EM step; coefficients at the old X are in co, every update uses the old X and Y.
// Control-variate and plain steps (Integrators/FKACIntegrator.hpp:46-57, :20-29; the plain
// step's g is 0): xi += Y v . dW with v = -sigma^T g.  VR step (FKACIntegrator.hpp:31-44; g is
// grad u / u, so v is mu): Y += Y mu . dW and the drift of X loses sigma mu; xi is not used.
// VR+CV step (FKACIntegrator.hpp:100-112, mu and F from the problem's table): xi += Y F . dW,
// Y += c Y h + Y mu . dW and the drift of X loses sigma mu.
// LT: the step carries local-time terms lz = psi Y L and ly = varphi Y L of a reflecting side
// (Integrators/FKACIntegrator.hpp:156-158: Z += f Y h + psi Y L, Y += c Y h + varphi Y L).
template <class R, int FAM, bool LT = false>
__device__ __forceinline__ void em_step(PathState<R> &s, const Coef<R> &co, R h, R sqrth, R n1, R n2, int fam,
                                        R lz = (R)0, R ly = (R)0) {
    R i0 = sqrth * n1, i1 = sqrth * n2;
    const bool vrcv = FAM != STEP_CV && fam == STEP_VRCV;
    R v0 = -(co.s00 * co.g0 + co.s10 * co.g1);
    R v1 = -(co.s01 * co.g0 + co.s11 * co.g1);
    if (vrcv) { v0 = co.F0; v1 = co.F1; }
    R dxi = s.Y * (v0 * i0 + v1 * i1);
    R dz = co.f * s.Y * h;
    if constexpr (LT) dz = dz + lz;
    const bool vr = FAM != STEP_CV && fam == STEP_VR;
    if constexpr (sizeof(R) == 8) {
        if (!vr) s.xi = s.xi + dxi;
        s.Z = s.Z + dz;
    } else {
        // Kahan-compensated functionals
        if (!vr) { R yk = dxi - s.xc; R tk = s.xi + yk; s.xc = (tk - s.xi) - yk; s.xi = tk; }
        R yz = dz - s.zc; R tz = s.Z + yz; s.zc = (tz - s.Z) - yz; s.Z = tz;
    }
    R Yold = s.Y;
    R b0 = co.b0, b1 = co.b1;
    if (vr || vrcv) {
        // mu: v for VR, the table's for VR+CV
        const R m0 = vrcv ? co.m0 : v0, m1 = vrcv ? co.m1 : v1;
        R dy = co.c * Yold * h + (vrcv ? Yold * (m0 * i0 + m1 * i1) : dxi);
        if constexpr (LT) dy = dy + ly;
        s.Y = Yold + dy;
        b0 = b0 - (co.s00 * m0 + co.s01 * m1);
        b1 = b1 - (co.s10 * m0 + co.s11 * m1);
    } else {
        R dy = co.c * Yold * h;
        if constexpr (LT) dy = dy + ly;
        s.Y = Yold + dy;
    }
    R nx = s.x + (b0 * h + (co.s00 * i0 + co.s01 * i1));
    R ny = s.y + (b1 * h + (co.s10 * i0 + co.s11 * i1));
    s.x = nx;
    s.y = ny;
}

// Normal pair of step s.st for this lane's path.  FP32: one Philox call per even step gives
// the normals of steps st and st + 1; `even` is the parity of st, which the path kernels keep
// the same for every lane of a wave (paths only start on odd wave iterations), so the draw is a
// wave-uniform branch taken every other step.
template <class R>
__device__ __forceinline__ void draw_normals(PathState<R> &s, uint32_t seed, uint32_t stream, uint32_t knot,
                                             R &n1, R &n2, bool even) {
    if constexpr (sizeof(R) == 8) {
        (void)even;
        double a, b;
        normals_f64(seed, stream, knot, s.path, s.st, a, b);
        n1 = a; n2 = b;
    } else {
        if (even) {
            const uint32_t call = s.st >> 1;
            Philox4 w = philox4x32_10(call, s.path, knot, 0u, seed, stream);
            // u = (k + 1/2) 2^-24 as one fma (one rounding of the same exact value as k + 1/2);
            // -2 ln u = (-2 ln 2) log2 u with the hardware log2
            const float sc = 5.9604644775390625e-08f, half = 2.98023223876953125e-08f;
            float u1 = __builtin_fmaf((float)(w.v[0] >> 8), sc, half);
            float u2 = __builtin_fmaf((float)(w.v[1] >> 8), sc, half);
            float u3 = __builtin_fmaf((float)(w.v[2] >> 8), sc, half);
            float u4 = __builtin_fmaf((float)(w.v[3] >> 8), sc, half);
            float r1 = RM<float>::sqrt(__builtin_amdgcn_logf(u1) * -1.3862943611198906f);
            float r2 = RM<float>::sqrt(__builtin_amdgcn_logf(u3) * -1.3862943611198906f);
            n1 = r1 * __builtin_amdgcn_cosf(u2);
            n2 = r1 * __builtin_amdgcn_sinf(u2);
            s.pn1 = r2 * __builtin_amdgcn_cosf(u4);
            s.pn2 = r2 * __builtin_amdgcn_sinf(u4);
        } else {
            n1 = s.pn1;
            n2 = s.pn2;
        }
    }
}

// ----------------------------------------------------------------------------
// Reflecting sides (Neumann / Robin): the Lepingle step (Integrators/FKACIntegrator.hpp:114-165)
// ----------------------------------------------------------------------------
// Whether the nearest side `kind` of a path (rect_dist's kind) reflects: bit s of rmask marks a
// reflecting side; a corner reflects only if both of its sides do.
PDDS_HD bool refl_kind(int kind, int rmask) {
    if (kind < 4) return ((rmask >> kind) & 1) != 0;
    const int v = (kind == 4 || kind == 5) ? 0 : 1;     // N for NE/NW, S for SE/SW
    const int hs = (kind == 4 || kind == 6) ? 2 : 3;    // E for NE/SE, W for NW/SW
    return ((rmask >> v) & 1) && ((rmask >> hs) & 1);
}

// Exponential variate (mean 1) of step st for the bridge maximum: a Philox draw of its own
// (counter word 3 with its top bit set, so it never meets a normal draw's counter).
template <class R>
__device__ __forceinline__ R draw_exponential(uint32_t st, uint32_t path, uint32_t seed, uint32_t stream,
                                              uint32_t knot) {
    if constexpr (sizeof(R) == 8) {
        Philox4 w = philox4x32_10(st, path, knot, 0x80000000u, seed, stream);
        return -dlog_core(u01_53(w.v[0], w.v[1]), 0);
    } else {
        Philox4 w = philox4x32_10(st, path, knot, 0x80000000u, seed, stream);
        const float u = __builtin_fmaf((float)(w.v[0] >> 8), 5.9604644775390625e-08f, 2.98023223876953125e-08f);
        return __builtin_amdgcn_logf(u) * -0.6931471805599453f;
    }
}

// One step of a path whose nearest side (outward normal (nx, ny), signed distance d <= 0)
// reflects.  lep: the Lepingle step -- the local time L is how far the maximum of the Brownian
// bridge of the step in the normal direction, M = (u + sqrt(u^2 + 2 |N^T sigma|^2 h E)) / 2 with
// u the step's normal displacement and E ~ Exp(1), passes the side, L = max(0, M + d), and the
// new X is pushed back by L along the normal; else a plain step (L = 0).  The functionals take
// this step's L and the local time lt of a projection at the end of the previous step
// (Integrators/FKACIntegrator.hpp:156-158), with psi at the old X's projection onto the side
// (the reference takes it at the old X, O(sqrt h) inside, which costs half an order).
template <class R, int PDE, int FAM>
__device__ __forceinline__ void refl_step(PathState<R> &s, const Coef<R> &co, const PdeParams &pp, R h, R sqrth,
                                          R n1, R n2, int fam, bool lep, R E, R nx, R ny, R d, R lt) {
    R L = (R)0;
    if (lep) {
        const R i0 = sqrth * n1, i1 = sqrth * n2;
        const R ga = nx * co.s00 + ny * co.s10, gb = nx * co.s01 + ny * co.s11;
        const R uc = (ga * i0 + gb * i1) + (nx * co.b0 + ny * co.b1) * h;
        const R M = (R)0.5 * (uc + RM<R>::sqrt((ga * ga + gb * gb) * (((R)2.0 * h) * E) + uc * uc));
        L = RM<R>::fmax((R)0, M + d);
    }
    const R Lt = L + lt;
    R lz = (R)0, ly = (R)0;
    if (Lt > (R)0) {
        // boundary data at X's projection onto the side, X - d n; the local time weighed by
        // (e^{varphi L} - 1) / (varphi L), Y's exact growth over it (1 on a Neumann side)
        const R phi = (R)pp.robin;
        const R psi = pde_psi<PDE, R>(s.x - d * nx, s.y - d * ny, nx, ny, phi, pp);
        const R w = Lt * RM<R>::exprel(phi * Lt);
        lz = (psi * s.Y) * w;
        ly = (phi * s.Y) * w;
    }
    em_step<R, FAM, true>(s, co, h, sqrth, n1, n2, fam, lz, ly);
    if (lep) {
        s.x = s.x - L * nx;
        s.y = s.y - L * ny;
    }
}

// Wave iterations on which a lane may start a path: FP32 paths start only on odd iterations,
// so that a stepping lane's step count has the parity of the iteration; FP64 paths any time.
template <class R> __device__ __forceinline__ bool may_start(int it) { return sizeof(R) == 8 || (it & 1); }

// Lane 0's value in every lane, held in scalar registers: the queue bookkeeping of a wave
// (pool bounds, queue index) is uniform and does not need vector registers.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Knot of global path g: the last knot whose first path is <= g.
__device__ __forceinline__ int find_knot(const KnotDesc *knots, int n_knots, int64_t g) {
    int lo = 0, hi = n_knots;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (knots[mid].path_off <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Exit record of a stopped path at batch slot g: the exit projection, Y and the score
// (Z + xi, plus Y g(X) for a global-boundary exit, Reduce_Analytic
// Solvers/EMFKACSolver.hpp:559-566); FP32 paths also put their exit side in Y's low bits.
template <class R, int PDE, int GEOM>
__device__ __forceinline__ void write_exit(const PathKernelArgs &a, uint32_t g, const PathState<R> &s,
                                           const R *dom, int gmask, R ex, R ey, int kind) {
    int side;
    if constexpr (sizeof(R) == 8) {
        if constexpr (GEOM == GeomDisc) side = classify_sector(dom, a.glob[2], (double)ex, (double)ey);
        else side = classify_coord(dom, a.glob, (double)ex, (double)ey);
    } else {
        side = classify_kind(kind, gmask);
    }
    double sc;
    if (side == 4) {
        R gv = pde_g<PDE, R>(ex, ey, a.pp);
        sc = (double)(s.Z + s.Y * gv + s.xi);
    } else {
        sc = (double)(s.Z + s.xi);
    }
    ExitRec rc;
    rc.x = (double)ex;
    rc.y = (double)ey;
    rc.Y = (double)s.Y;
    if constexpr (sizeof(R) == 4)
        rc.Y = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, rc.Y) | (uint64_t)(side + 1));
    rc.score = sc;
    a.rec.r[g] = rc;
}

// Steps and step-cap stops of stopped paths of one knot, summed in a lane's registers and added
// to the knot's counters (Records::kacc) when the lane moves to another knot and when it ends.
struct KnotSteps {
    int k;                    // knot of the sums
    unsigned long long st;    // steps
    uint32_t cap;             // paths stopped by the step cap while inside

    __device__ __forceinline__ void init() { k = 0; st = 0ULL; cap = 0u; }
    __device__ __forceinline__ void flush(const PathKernelArgs &a) {
        if (st) {
            atomicAdd(a.rec.kacc + 2 * (int64_t)k, st);
            if (cap) atomicAdd(a.rec.kacc + 2 * (int64_t)k + 1, (unsigned long long)cap);
            st = 0ULL;
            cap = 0u;
        }
    }
    // the lane takes a path of knot kn
    __device__ __forceinline__ void enter(const PathKernelArgs &a, int kn) {
        if (kn != k) {
            flush(a);
            k = kn;
        }
    }
};

// Distance to a stencil's stopping domain: its rectangle, or its annular sector on a disc mesh.
template <class R, int GEOM>
__device__ __forceinline__ R domain_dist(const R *dom, R x, R y, R &ex, R &ey, R &nx, R &ny, int &kind) {
    if constexpr (GEOM == GeomDisc) return sector_dist<R>(dom, x, y, ex, ey, nx, ny, kind);
    else return rect_dist<R>(dom, x, y, ex, ey, nx, ny, kind);
}
template <int GEOM> __device__ __forceinline__ const double *domain_of(const StencilDesc &sd) {
    if constexpr (GEOM == GeomDisc) return sd.geo;
    else return sd.rect;
}
template <int GEOM> constexpr int domain_len() { return GEOM == GeomDisc ? 6 : 4; }

// ----------------------------------------------------------------------------
// Per-lane walker: the state of the path a lane carries and one step of it
// ----------------------------------------------------------------------------
// MIXED: the knots of the launch differ in Philox stream, Helmholtz k2 (serving rounds) or step h
// (adaptive h), so each lane carries its knot's; otherwise they are the launch's and stay in
// scalar registers.
// REFL: some global sides reflect (rectangles only): the path carries its nearest side (distance,
// kind, outward normal) and the local time of a projection, and steps with refl_step while that
// side reflects; only absorbing sides stop it.
template <class R, int PDE, bool MIXED, int GEOM, int FAM, bool REFL = false>
struct Walker {
    PathState<R> s;
    Coef<R> co;            // coefficients at the current X
    R pn1, pn2;            // normals of step s.st (the next step)
    R rect[domain_len<GEOM>()];   // stopping domain in path precision (FP64 classifies on it)
    uint32_t kid;
    int gmask;
    uint32_t g;            // batch path index (record slot; a batch holds fewer than 2^32 paths)
    uint32_t kstream;      // Philox key word 1 of the path's knot
    R k2;                  // Helmholtz parameter of the path's knot
    R hk, sqhk;            // step of the path's knot and its square root
    bool active, fresh;
    // REFL: reflecting sides of the stopping domain, nearest side at the current X, pending
    // local time
    int rmask, rk;
    R rd, rnx, rny, lt;
    KnotSteps ks;          // steps of the lane's stopped paths not yet added to their knot's counters

    __device__ __forceinline__ void init() {
        pn1 = 0; pn2 = 0; kid = 0; kstream = 0; gmask = 0; g = 0; k2 = 0; hk = 0; sqhk = 0;
        active = false; fresh = false;
        rmask = 0; rk = 0; rd = 0; rnx = 0; rny = 0; lt = 0;
        ks.init();
    }

    // take path gg of the batch; k is its knot
    __device__ __forceinline__ void start(const PathKernelArgs &a, int64_t gg, int k) {
        ks.enter(a, k);
        const KnotDesc &kd = a.knots[k];
        const StencilDesc &sd = a.tables[kd.table].sten[kd.stencil];
        const double *dom = domain_of<GEOM>(sd);
        for (int q = 0; q < domain_len<GEOM>(); ++q) rect[q] = (R)dom[q];
        gmask = sd.gmask;
        kid = kd.knot_id;
        if constexpr (MIXED) {
            kstream = kd.stream;
            k2 = (R)kd.k2;
            hk = (R)kd.h;
            sqhk = (R)kd.sqrth;
        }
        if constexpr (REFL) {
            rmask = sd.gmask & a.pp.refl;   // stencil sides on reflecting global sides
            lt = (R)0;
        }
        g = (uint32_t)gg;
        s.path = (uint32_t)(gg - kd.path_off);
        s.x = (R)kd.x0; s.y = (R)kd.y0;
        s.Y = (R)1; s.Z = (R)0; s.xi = (R)0; s.zc = (R)0; s.xc = (R)0;
        s.st = 0;
        active = true;
        fresh = true;
    }

    // One step for every lane, as one straight-line block (a lane that is not stepping
    // computes and discards), so the LUT loads of the new X are issued together and overlap
    // the next step's normal draw, which does not depend on X.  Inactive lanes hold stale or
    // unset state: they touch no memory except the LUT, whose coordinates are clamped into
    // the table.  A lane whose path stops writes its exit record and goes inactive.
    // DRAW_FIRST: draw this step's normals at its start instead of overlapping the next
    // step's draw with the LUT reads (for LDS-resident LUT tiles, whose reads are short).
    // it: the wave's iteration count (stepping lanes have s.st of its parity)
    template <bool TILED, bool DRAW_FIRST>
    __device__ __forceinline__ void step(const PathKernelArgs &a, const LutTile &T, int it) {
        const R h = MIXED ? hk : (R)a.h, sqrth = MIXED ? sqhk : (R)a.sqrth, beta = (R)a.beta;
        const uint32_t strm = MIXED ? kstream : a.stream;
        const R kk2 = MIXED ? k2 : (R)a.pp.k2;
        const bool stepping = active && !fresh;
        if constexpr (DRAW_FIRST) draw_normals<R>(s, a.seed, strm, kid, pn1, pn2, (it & 1) == 0);
        {
            PathState<R> t = s;
            if constexpr (REFL) {
                // the nearest side at the current X reflects: the Lepingle step within rho of it
                const bool rside = refl_kind(rk, rmask);
                const bool lep = rside && (a.pp.rho <= 0.0 || rd > -(R)a.pp.rho);
                R E = (R)0;
                if (stepping && lep) E = draw_exponential<R>(s.st, s.path, a.seed, strm, kid);
                refl_step<R, PDE, FAM>(t, co, a.pp, h, sqrth, pn1, pn2, a.pp.step, lep, E, rnx, rny, rd, lt);
            } else {
                em_step<R, FAM>(t, co, h, sqrth, pn1, pn2, a.pp.step);
            }
            s.x = stepping ? t.x : s.x;
            s.y = stepping ? t.y : s.y;
            s.Y = stepping ? t.Y : s.Y;
            s.Z = stepping ? t.Z : s.Z;
            s.xi = stepping ? t.xi : s.xi;
            if constexpr (sizeof(R) == 4) {
                s.zc = stepping ? t.zc : s.zc;
                s.xc = stepping ? t.xc : s.xc;
            }
            s.st += stepping ? 1 : 0;
        }
        R ex, ey, nx, ny;
        int kind = 0;
        // the stopping test needs only the distance and the nearest side's normal; on a rectangle
        // without reflecting sides those come without branches, and the exit projection and side
        // are formed below for the stopping lanes alone
        constexpr bool fast = GEOM == GeomRect && !REFL;
        R d;
        if constexpr (fast) d = rect_dist_inside<R>(rect, s.x, s.y, nx, ny);
        else d = domain_dist<R, GEOM>(rect, s.x, s.y, ex, ey, nx, ny, kind);
        bool rnow = false;
        if constexpr (REFL) {
            // past a reflecting side: back onto it, the distance becoming local time for the next
            // step (Inside's reflect outcome, Integrators/FKACIntegrator.hpp:189-207)
            rnow = refl_kind(kind, rmask);
            lt = (R)0;
            if (stepping && rnow && d > (R)0) {
                lt = d;
                s.x = ex;
                s.y = ey;
                d = domain_dist<R, GEOM>(rect, s.x, s.y, ex, ey, nx, ny, kind);
                rnow = refl_kind(kind, rmask);
            }
            rd = d; rk = kind; rnx = nx; rny = ny;
        }
        pde_coef_t<PDE, R, TILED, FAM>(s.x, s.y, a.pp, T, kk2, co);
        if constexpr (!DRAW_FIRST) draw_normals<R>(s, a.seed, strm, kid, pn1, pn2, (it & 1) != 0);
        const R ga = nx * co.s00 + ny * co.s10;
        const R gb = nx * co.s01 + ny * co.s11;
        const R nrm = RM<R>::sqrt(ga * ga + gb * gb);
        // a reflecting nearest side never stops the path
        const bool go = rnow || d < beta * nrm * sqrth;
        const bool capped = (int64_t)s.st >= a.max_steps;
        if (stepping && (!go || capped)) {
            if constexpr (fast) {
                R tx, ty;
                (void)rect_dist<R>(rect, s.x, s.y, ex, ey, tx, ty, kind);
            }
            write_exit<R, PDE, GEOM>(a, g, s, rect, gmask, ex, ey, kind);
            ks.st += s.st;
            ks.cap += capped && go ? 1u : 0u;
            active = false;
        }
        fresh = false;
    }
};

// Wave-private pool of consecutive paths [next, end) taken a.pool_chunk at a time from one
// queue (one atomic per a.pool_chunk paths: 64 lanes x the run's paths per lane).  With it, the knot of the pool's first path and
// where that knot ends, found once per chunk with wave-uniform values; a lane whose path
// lies past that knot searches for its own.
struct WavePool {
    int64_t next = 0, end = 0;
    int ck = 0;
    int64_t ck_end = 0;

    // a fresh chunk from counter over [qb, qe); false if that queue is drained
    __device__ __forceinline__ bool take(const PathKernelArgs &a, unsigned long long *counter, int64_t qb,
                                         int64_t qe, int lane) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(counter, (unsigned long long)a.pool_chunk);
        base = uniform_u64(base);
        const int64_t s0 = qb + (int64_t)base;
        if (s0 >= qe) return false;
        next = s0;
        end = s0 + a.pool_chunk < qe ? s0 + a.pool_chunk : qe;
        ck = find_knot(a.knots, a.n_knots, s0);
        ck_end = ck + 1 < a.n_knots ? a.knots[ck + 1].path_off : a.total;
        return true;
    }
    __device__ __forceinline__ int knot_of(const PathKernelArgs &a, int64_t gg) const {
        return gg < ck_end ? ck : find_knot(a.knots, a.n_knots, gg);
    }
    // start paths on the lanes that need one; lanes left without a path once `drained`
    // is set are marked done
    template <class W>
    __device__ __forceinline__ void hand_out(const PathKernelArgs &a, W &w, bool need, unsigned long long m,
                                             bool drained, bool &done, int lane) {
        const int64_t avail = end > next ? end - next : 0;
        const int rank = __popcll(m & ((1ULL << lane) - 1ULL));
        if (need) {
            if (rank < avail) {
                const int64_t gg = next + rank;
                w.start(a, gg, knot_of(a, gg));
            } else if (drained) {
                done = true;
            }
        }
        const int64_t cnt = __popcll(m);
        next += cnt < avail ? cnt : avail;
    }
};

// ----------------------------------------------------------------------------
// Persistent path kernel: one path per lane, lanes refilled from XCD-local queues
// ----------------------------------------------------------------------------
// the run-time step-family variants carry the VR branch: one wave less (where more than two),
// so that they do not spill
template <class R, int PDE, int FAM> constexpr int path_waves_f() {
    return FAM == STEP_CV || path_waves<R, PDE>() < 3 ? path_waves<R, PDE>() : path_waves<R, PDE>() - 1;
}

template <class R, int PDE, bool MIXED, int GEOM, int FAM, bool REFL = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(path_waves_f<R, PDE, FAM>()))) path_kernel(PathKernelArgs a) {
    const int lane = __lane_id();
    const LutTile none{nullptr, 0, 0, 0, 0};
    Walker<R, PDE, MIXED, GEOM, FAM, REFL> w;
    w.init();
    bool done = false;
    // The wave starts on its XCD's queue, so the paths (and LUT texels) of one XCD's L2
    // stay in one contiguous range of knots; a drained queue sends the wave on to the next.
    WavePool pool;
    bool drained = false;
    int q = (int)(blockIdx.x % (unsigned)a.n_queues);
    int tried = 0;
    for (int it = 1;; ++it) {
        const bool can_start = may_start<R>(it);
        if (can_start) {
            const bool need = !w.active && !done;
            const unsigned long long m = __ballot(need);
            if (m) {
                while (pool.next >= pool.end && !drained) {
                    if (pool.take(a, a.counter + q, a.total * q / a.n_queues, a.total * (q + 1) / a.n_queues, lane)) break;
                    if (++tried >= a.n_queues) drained = true;
                    else q = q + 1 == a.n_queues ? 0 : q + 1;
                }
                pool.hand_out(a, w, need, m, drained, done, lane);
            }
        }
        if (!__any(w.active)) {
            if (can_start) break;
            continue;   // idle until the next iteration on which paths may start
        }
        w.template step<false, false>(a, none, it);
    }
    w.ks.flush(a);
}

// ----------------------------------------------------------------------------
// Plain path kernel, the baseline GPU form (the shape of the reference's CUDA SolveLoop without
// its chunked relaunches, Solvers/SolveLoop.cuh:20-887): one work-item per path of the batch,
// each stepping its own path to the exit; a wave lasts as long as its slowest path.
// ----------------------------------------------------------------------------
// FP64 control-variate plain kernels fit three waves per SIMD without spilling (on cfg2 that runs
// them 23 % faster than the compiler's own two-wave allocation); the run-time family variants
// would spill there and take two.  No kernel is given a one-wave budget: above 256 registers the
// compiler spills into accumulation registers, and the reflecting FP64 plain kernel built so
// (264 registers, 8 of them accumulation registers) returned wrong G rows while its two-wave build
// (256 registers) is bit-exact with the oracle.
template <class R, int FAM> constexpr int plain_waves() { return sizeof(R) == 8 && FAM == STEP_CV ? 3 : 2; }

template <class R, int PDE, int GEOM, int FAM, bool REFL = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(plain_waves<R, FAM>()))) plain_kernel(PathKernelArgs a) {
    const int64_t gg = a.base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gg >= a.total) return;
    const int kn = find_knot(a.knots, a.n_knots, gg);
    const KnotDesc &kd = a.knots[kn];
    const StencilDesc &sd = a.tables[kd.table].sten[kd.stencil];
    const LutTile none{nullptr, 0, 0, 0, 0};
    const R h = (R)kd.h, sqrth = (R)kd.sqrth, beta = (R)a.beta, k2 = (R)kd.k2;
    R rect[domain_len<GEOM>()];
    const double *dom = domain_of<GEOM>(sd);
    for (int q = 0; q < domain_len<GEOM>(); ++q) rect[q] = (R)dom[q];
    PathState<R> s;
    s.path = (uint32_t)(gg - kd.path_off);
    s.x = (R)kd.x0; s.y = (R)kd.y0;
    s.Y = (R)1; s.Z = (R)0; s.xi = (R)0; s.zc = (R)0; s.xc = (R)0;
    s.st = 0;
    Coef<R> co;
    pde_coef_t<PDE, R, false, FAM>(s.x, s.y, a.pp, none, k2, co);
    R ex, ey;
    int kind;
    // REFL: reflecting sides, the nearest side at the current X and the pending local time (as
    // in Walker::step)
    const int rmask = REFL ? (sd.gmask & a.pp.refl) : 0;
    R rd = (R)0, rnx = (R)0, rny = (R)0, lt = (R)0;
    int rk = 0;
    if constexpr (REFL) rd = domain_dist<R, GEOM>(rect, s.x, s.y, ex, ey, rnx, rny, rk);
    bool go, capped;
    do {
        R n1, n2, nx, ny;
        draw_normals<R>(s, a.seed, kd.stream, kd.knot_id, n1, n2, (s.st & 1) == 0);
        if constexpr (REFL) {
            const bool lep = refl_kind(rk, rmask) && (a.pp.rho <= 0.0 || rd > -(R)a.pp.rho);
            const R E = lep ? draw_exponential<R>(s.st, s.path, a.seed, kd.stream, kd.knot_id) : (R)0;
            refl_step<R, PDE, FAM>(s, co, a.pp, h, sqrth, n1, n2, a.pp.step, lep, E, rnx, rny, rd, lt);
        } else {
            em_step<R, FAM>(s, co, h, sqrth, n1, n2, a.pp.step);
        }
        s.st += 1;
        R d = domain_dist<R, GEOM>(rect, s.x, s.y, ex, ey, nx, ny, kind);
        bool rnow = false;
        if constexpr (REFL) {
            rnow = refl_kind(kind, rmask);
            lt = (R)0;
            if (rnow && d > (R)0) {
                lt = d;
                s.x = ex;
                s.y = ey;
                d = domain_dist<R, GEOM>(rect, s.x, s.y, ex, ey, nx, ny, kind);
                rnow = refl_kind(kind, rmask);
            }
            rd = d; rk = kind; rnx = nx; rny = ny;
        }
        pde_coef_t<PDE, R, false, FAM>(s.x, s.y, a.pp, none, k2, co);
        const R ga = nx * co.s00 + ny * co.s10;
        const R gb = nx * co.s01 + ny * co.s11;
        const R nrm = RM<R>::sqrt(ga * ga + gb * gb);
        go = rnow || d < beta * nrm * sqrth;
        capped = (int64_t)s.st >= a.max_steps;
    } while (go && !capped);
    write_exit<R, PDE, GEOM>(a, (uint32_t)gg, s, rect, sd.gmask, ex, ey, kind);
    // steps and step-cap stops into the knot's counters: one wave sum and one atomic per counter
    // when all 64 lanes hold paths of one knot, else an atomic per lane
    const bool cap = capped && go;
    const int k0 = __builtin_amdgcn_readfirstlane(kn);
    if (__ballot(1) == ~0ULL && __all(kn == k0)) {
        unsigned long long v = s.st;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        const unsigned long long nc = (unsigned long long)__popcll(__ballot(cap));
        if (__lane_id() == 0) {
            atomicAdd(a.rec.kacc + 2 * (int64_t)kn, v);
            if (nc) atomicAdd(a.rec.kacc + 2 * (int64_t)kn + 1, nc);
        }
    } else {
        atomicAdd(a.rec.kacc + 2 * (int64_t)kn, (unsigned long long)s.st);
        if (cap) atomicAdd(a.rec.kacc + 2 * (int64_t)kn + 1, 1ULL);
    }
}

// ----------------------------------------------------------------------------
// LDS-tiled path kernel (coefficient-LUT problem): one work-group per CU.  The work-group
// works on one patch at a time: it copies the patch's plane-a tile into LDS, its waves take
// paths from the patch's queue until it is drained, and once every wave is done it moves to
// the next patch with paths left.  The 16 taps of a step are then LDS reads; only field 4
// (one 16-byte load per tap row, 4 per step, from its float plane) and taps outside the tile
// go to the vector L1.
// ----------------------------------------------------------------------------
// one work-group per CU holding every wave the register budget allows (the run-time step-family
// variants one wave less, as in path_kernel, and so the FP32 annular-sector kernels, whose
// sector distance needs more registers than three waves per SIMD leave)
template <class R, int FAM, int GEOM> constexpr int tiled_waves() {
    return path_waves_f<R, PDE_VARCOEF_LUT, FAM>() -
           (sizeof(R) == 4 && GEOM == GeomDisc && path_waves_f<R, PDE_VARCOEF_LUT, FAM>() > 2 ? 1 : 0);
}
template <class R, int FAM, int GEOM> constexpr int tiled_wg() { return 64 * 4 * tiled_waves<R, FAM, GEOM>(); }

__device__ __forceinline__ int find_patch(const PatchDesc *p, int n, int64_t g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (p[mid].pbeg <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class R, bool MIXED, int GEOM, int FAM>
__global__ void __launch_bounds__((tiled_wg<R, FAM, GEOM>())) __attribute__((amdgpu_waves_per_eu(tiled_waves<R, FAM, GEOM>())))
path_kernel_tiled(PathKernelArgs a) {
    extern __shared__ float4 tile[];
    __shared__ int s_patch;
    const int lane = __lane_id();
    Walker<R, PDE_VARCOEF_LUT, MIXED, GEOM, FAM> w;
    w.init();
    // first patch: the one holding this work-group's share of the batch's path range
    int p = find_patch(a.patches, a.n_patches, (int64_t)blockIdx.x * a.total / (int64_t)gridDim.x);
    while (true) {
        if (threadIdx.x == 0) {
            int pick = -1;
            for (int k = 0; k < a.n_patches; ++k) {
                const int c = p + k < a.n_patches ? p + k : p + k - a.n_patches;
                const PatchDesc &pd = a.patches[c];
                const unsigned long long taken = __hip_atomic_load(a.counter + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((int64_t)taken < pd.pend - pd.pbeg) { pick = c; break; }
            }
            s_patch = pick;
        }
        __syncthreads();
        p = __builtin_amdgcn_readfirstlane(s_patch);   // uniform: the patch, its tile and queue stay scalar
        if (p < 0) break;
        const PatchDesc pd = a.patches[p];
        const LutTile T{(lds_float4 *)tile, pd.c0, pd.r0, pd.tw, pd.tiled ? pd.th : 0, pd.pitch};
        if (pd.tiled) {
            const int nt = pd.tw * pd.th;
            for (int i = threadIdx.x; i < nt; i += blockDim.x) {
                const int r = i / pd.tw, c = i - r * pd.tw;
                tile[r * pd.pitch + c] = a.pp.lut.a[(int64_t)(pd.r0 + r) * a.pp.lut.px + (pd.c0 + c)];
            }
        }
        __syncthreads();
        WavePool pool;
        bool drained = false, done = false;
        for (int it = 1;; ++it) {
            const bool can_start = may_start<R>(it);
            if (can_start) {
                const bool need = !w.active && !done;
                const unsigned long long m = __ballot(need);
                if (m) {
                    if (pool.next >= pool.end && !drained)
                        if (!pool.take(a, a.counter + p, pd.pbeg, pd.pend, lane)) drained = true;
                    pool.hand_out(a, w, need, m, drained, done, lane);
                }
            }
            if (!__any(w.active)) {
                if (can_start) break;
                continue;
            }
            // FP64 overlaps the next step's normal draw with the LUT reads; the shorter FP32
            // draw goes first (measured: each order is about 1 % faster in its precision)
            w.template step<true, sizeof(R) == 4>(a, T, it);
        }
        __syncthreads();   // every wave is done with this tile
    }
    w.ks.flush(a);
}

// ----------------------------------------------------------------------------
// Reduction kernel: one work-group per work item
// ----------------------------------------------------------------------------
template <bool EQ, int SIDE> __global__ void __launch_bounds__(kReduceWG) reduce_kernel(ReduceArgs a);

// One block per knot: sum the knot's items in item order.
__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f);

// ----------------------------------------------------------------------------
// Per-path kernel (MCinCUDA shape): one path per lane, records written out.
// ----------------------------------------------------------------------------
// REFL: the sides of a.rect in the mask a.pp.refl reflect (as in Walker::step).
template <class R, int PDE, int FAM, bool REFL = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) paths_kernel(PathsArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const R h = (R)a.h, sqrth = (R)a.sqrth;
    const R rect[4] = {(R)a.rect[0], (R)a.rect[1], (R)a.rect[2], (R)a.rect[3]};
    PathState<R> s;
    s.path = (uint32_t)(a.path0 + i);
    s.x = (R)a.x0; s.y = (R)a.y0;
    s.Y = (R)1; s.Z = (R)0; s.xi = (R)0; s.zc = (R)0; s.xc = (R)0;
    s.st = 0;
    Coef<R> co;
    pde_coef<PDE, R, FAM>(s.x, s.y, a.pp, co);
    R ex = s.x, ey = s.y;
    const int rmask = REFL ? (a.pp.refl & 15) : 0;
    R rd = (R)0, rnx = (R)0, rny = (R)0, lt = (R)0;
    int rk = 0;
    if constexpr (REFL) rd = rect_dist<R>(rect, s.x, s.y, ex, ey, rnx, rny, rk);
    for (int phase = 0; phase < (a.phase2 ? 2 : 1); ++phase) {
        const R beta = phase == 0 ? (R)a.beta : (R)0;
        const uint32_t st0 = s.st;
        bool go = true;
        while (go) {
            R n1, n2, nx, ny;
            int kind;
            draw_normals<R>(s, a.seed, a.stream, a.knot, n1, n2, (s.st & 1) == 0);
            if constexpr (REFL) {
                const bool lep = refl_kind(rk, rmask) && (a.pp.rho <= 0.0 || rd > -(R)a.pp.rho);
                const R E = lep ? draw_exponential<R>(s.st, s.path, a.seed, a.stream, a.knot) : (R)0;
                refl_step<R, PDE, FAM>(s, co, a.pp, h, sqrth, n1, n2, a.pp.step, lep, E, rnx, rny, rd, lt);
            } else {
                em_step<R, FAM>(s, co, h, sqrth, n1, n2, a.pp.step);
            }
            s.st += 1;
            R d = rect_dist<R>(rect, s.x, s.y, ex, ey, nx, ny, kind);
            bool rnow = false;
            if constexpr (REFL) {
                rnow = refl_kind(kind, rmask);
                lt = (R)0;
                if (rnow && d > (R)0) {
                    lt = d;
                    s.x = ex;
                    s.y = ey;
                    d = rect_dist<R>(rect, s.x, s.y, ex, ey, nx, ny, kind);
                    rnow = refl_kind(kind, rmask);
                }
                rd = d; rk = kind; rnx = nx; rny = ny;
            }
            pde_coef<PDE, R, FAM>(s.x, s.y, a.pp, co);
            R ga = nx * co.s00 + ny * co.s10;
            R gb = nx * co.s01 + ny * co.s11;
            R nrm = RM<R>::sqrt(ga * ga + gb * gb);
            go = (rnow || d < beta * nrm * sqrth) && ((int64_t)(s.st - st0) < a.max_steps);
        }
        double *o[5];
        int64_t *os;
        if (phase == 0) { o[0] = a.xt; o[1] = a.yt; o[2] = a.Y; o[3] = a.Z; o[4] = a.xi; os = a.steps; }
        else { o[0] = a.xt2; o[1] = a.yt2; o[2] = a.Y2; o[3] = a.Z2; o[4] = a.xi2; os = a.steps2; }
        if (o[0]) o[0][i] = (double)ex;
        if (o[1]) o[1][i] = (double)ey;
        if (o[2]) o[2][i] = (double)s.Y;
        if (o[3]) o[3][i] = (double)s.Z;
        if (o[4]) o[4][i] = (double)s.xi;
        if (os) os[i] = (int64_t)(s.st - st0);
    }
}

}  // namespace pdds
