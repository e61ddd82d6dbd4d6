// engine_kernels.h -- HIP kernels of the path engine (templated on path precision R
// and on the PDE functor).  Included by engine_f64.hip (compiled with
// -ffp-contract=off, bit-exact with the FP64 oracle) and engine_f32.hip.
//
// Work decomposition: the paths of knot k are cut into work items of P paths;
// an item is run by one work-group.  Lanes take paths from an LDS counter and
// refill as soon as their path stops (lane recycling, the on-device form of the
// reference's 1000-step relaunch loop, Solvers/SolveLoop.cuh:59-74,
// solveCUDA.cuh:118-168).  In persistent mode work-groups take items from a
// device-wide atomic queue until it is empty.  Each path is stepped by the
// control-variate Euler-Maruyama scheme (Integrators/FKACIntegrator.hpp:46-57)
// and stopped by the Gobet-shifted test (Integrators/FKACIntegrator.hpp:168-186).
// When an item's paths are done, the work-group reduces them in place: per
// sub-batch of S paths the per-side RBF moments Phi_j = sum Y phi_j(X_tau)
// (m exps per exit instead of the reference's m^2, Meshes/stencil.cpp:244-307),
// then G_beta = -Psi^-T Phi per CSR entry, accumulated as sum and sum of squares
// (batch-means variance) into per-item partials; a finalize kernel sums the
// items of each knot in item order.  Every sum has a fixed order, so results do
// not depend on scheduling.
#pragma once
#include "pdds_pde.h"
#include "engine_types.h"

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

// Gaussian RBF exp(-|p - X|^2 / c^2), sqrt-then-square as BVPs/BVP.hpp:84-87
PDDS_HD double rbf(double px, double py, double x, double y, double c) {
    double dx = px - x, dy = py - y;
    double r = __builtin_sqrt(dx * dx + dy * dy);
    return dexp(-(r * r) / (c * c));
}

// ----------------------------------------------------------------------------
// Per-lane path state
// ----------------------------------------------------------------------------
template <class R> struct PathState {
    R x, y, Y, Z, xi;
    R zc, xc;          // Kahan compensations (FP32 only)
    R pn1, pn2;        // pending normals of the odd step (FP32 only)
    int64_t st;
    uint32_t path;
    int plocal;
};

// One step of the control-variate EM scheme; coefficients at the old X are in co.
template <class R>
__device__ __forceinline__ void em_step(PathState<R> &s, const Coef<R> &co, R h, R sqrth, R n1, R n2) {
    R i0 = sqrth * n1, i1 = sqrth * n2;
    R v0 = -(co.s00 * co.g0 + co.s10 * co.g1);
    R v1 = -(co.s01 * co.g0 + co.s11 * co.g1);
    R dxi = s.Y * (v0 * i0 + v1 * i1);
    R dz = co.f * s.Y * h;
    if constexpr (sizeof(R) == 8) {
        s.xi = s.xi + dxi;
        s.Z = s.Z + dz;
    } else {
        // Kahan-compensated functionals
        R yk = dxi - s.xc; R tk = s.xi + yk; s.xc = (tk - s.xi) - yk; s.xi = tk;
        R yz = dz - s.zc; R tz = s.Z + yz; s.zc = (tz - s.Z) - yz; s.Z = tz;
    }
    R Yold = s.Y;
    s.Y = Yold + co.c * Yold * h;
    R nx = s.x + (co.b0 * h + (co.s00 * i0 + co.s01 * i1));
    R ny = s.y + (co.b1 * h + (co.s10 * i0 + co.s11 * i1));
    s.x = nx;
    s.y = ny;
}

// Normal pair of step s.st for this lane's path.
template <class R>
__device__ __forceinline__ void draw_normals(PathState<R> &s, uint32_t seed, uint32_t stream, uint32_t knot,
                                             R &n1, R &n2) {
    if constexpr (sizeof(R) == 8) {
        double a, b;
        normals_f64(seed, stream, knot, s.path, s.st, a, b);
        n1 = a; n2 = b;
    } else {
        if ((s.st & 1) == 0) {
            int64_t call = s.st >> 1;
            Philox4 w = philox4x32_10((uint32_t)call, s.path, knot, (uint32_t)(call >> 32), seed, stream);
            const float sc = 5.9604644775390625e-08f;
            float u1 = ((float)(w.v[0] >> 8) + 0.5f) * sc;
            float u2 = ((float)(w.v[1] >> 8) + 0.5f) * sc;
            float u3 = ((float)(w.v[2] >> 8) + 0.5f) * sc;
            float u4 = ((float)(w.v[3] >> 8) + 0.5f) * sc;
            float r1 = __builtin_sqrtf(-2.0f * RM<float>::log(u1));
            float r2 = __builtin_sqrtf(-2.0f * RM<float>::log(u3));
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
// Assembly kernel
// ----------------------------------------------------------------------------
template <class R, int PDE, bool PERSIST>
__global__ void __launch_bounds__(256) assemble_kernel(EngineArgs a) {
    extern __shared__ double smem[];
    __shared__ int64_t s_item;
    __shared__ int s_next;
    __shared__ unsigned long long s_steps;
    __shared__ unsigned int s_cnt[8];
    __shared__ double s_bsb[kNBG], s_bbsb[kNBG];

    const int tid = threadIdx.x;
    const int WG = blockDim.x;
    const int64_t slot = blockIdx.x;
    double *rx = a.rec_x + slot * a.rec_stride;
    double *ry = a.rec_y + slot * a.rec_stride;
    double *rY = a.rec_Y + slot * a.rec_stride;
    double *rs = a.rec_score + slot * a.rec_stride;
    int8_t *rside = a.rec_side + slot * a.rec_stride;
    double *Phi = smem;
    double *accS = smem + kNBG * a.ne_cap;
    double *accQ = accS + a.nnz_cap;

    int64_t item = PERSIST ? -1 : a.item_base + (int64_t)blockIdx.x;
    while (true) {
        if (PERSIST) {
            if (tid == 0) s_item = (int64_t)atomicAdd((unsigned long long *)a.work_counter, 1ULL);
            __syncthreads();
            // the item is the same for the whole work-group: keep it (and every descriptor
            // loaded through it) in scalar registers
            const int64_t v = s_item;
            item = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                             (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v));
        }
        if (item >= a.n_items) break;
        const ItemDesc it = a.items[item];
        const KnotDesc kd = a.knots[it.knot];
        const StencilDesc &sd = a.sten[kd.stencil];
        if (tid == 0) {
            s_next = 0;
            s_steps = 0ULL;
            for (int q = 0; q < 8; ++q) s_cnt[q] = 0u;
        }
        __syncthreads();

        // ---------------- path phase ----------------
        {
            const R h = (R)a.h, sqrth = (R)a.sqrth, beta = (R)a.beta;
            const R rect[4] = {(R)sd.rect[0], (R)sd.rect[1], (R)sd.rect[2], (R)sd.rect[3]};
            PathState<R> s;
            Coef<R> co;
            bool active = false, exhausted = false;
            unsigned long long my_steps = 0ULL;
            while (true) {
                bool fresh = false;
                if (!active && !exhausted) {
                    int p = atomicAdd(&s_next, 1);
                    if (p < it.n) {
                        s.plocal = p;
                        s.path = (uint32_t)(it.p_begin + p);
                        s.x = (R)kd.x0; s.y = (R)kd.y0;
                        s.Y = (R)1; s.Z = (R)0; s.xi = (R)0; s.zc = (R)0; s.xc = (R)0;
                        s.st = 0;
                        active = true;
                        fresh = true;
                    } else {
                        exhausted = true;
                    }
                }
                if (!__any(active)) break;
                R d = 0, ex = 0, ey = 0, nx = 0, ny = 0;
                int kind = 0;
                if (active && !fresh) {
                    R n1, n2;
                    draw_normals<R>(s, a.seed, a.stream, kd.knot_id, n1, n2);
                    em_step<R>(s, co, h, sqrth, n1, n2);
                    s.st += 1;
                    d = rect_dist<R>(rect, s.x, s.y, ex, ey, nx, ny, kind);
                }
                if (active) pde_coef<PDE, R>(s.x, s.y, a.pp, co);
                if (active && !fresh) {
                    R ga = nx * co.s00 + ny * co.s10;
                    R gb = nx * co.s01 + ny * co.s11;
                    R nrm = RM<R>::sqrt(ga * ga + gb * gb);
                    bool go = d < beta * nrm * sqrth;
                    bool capped = s.st >= a.max_steps;
                    if (!go || capped) {
                        int side;
                        if constexpr (sizeof(R) == 8) side = classify_coord(sd.rect, a.glob, (double)ex, (double)ey);
                        else side = classify_kind(kind, sd.gmask);
                        double sc;
                        if (side == 4) {
                            R g = pde_g<PDE, R>(ex, ey, a.pp);
                            sc = (double)(s.Z + s.Y * g + s.xi);
                        } else {
                            sc = (double)(s.Z + s.xi);
                        }
                        rx[s.plocal] = (double)ex;
                        ry[s.plocal] = (double)ey;
                        rY[s.plocal] = (double)s.Y;
                        rs[s.plocal] = sc;
                        rside[s.plocal] = (int8_t)side;
                        atomicAdd(&s_cnt[side < 0 ? 5 : side], 1u);
                        if (capped && go) atomicAdd(&s_cnt[6], 1u);
                        my_steps += (unsigned long long)s.st;
                        active = false;
                    }
                }
            }
            atomicAdd(&s_steps, my_steps);
        }
        __syncthreads();

        // ---------------- reduction phase ----------------
        const int ne = sd.msz[0] + sd.msz[1] + sd.msz[2] + sd.msz[3];
        const int nnz = sd.csr_nnz;
        const int S = a.S;
        const int nsb = (it.n + S - 1) / S;
        int eoff[5];
        eoff[0] = 0;
        for (int q = 0; q < 4; ++q) eoff[q + 1] = eoff[q] + sd.msz[q];
        for (int c = tid; c < nnz; c += WG) { accS[c] = 0.0; accQ[c] = 0.0; }
        double Bs = 0.0, BBs = 0.0;
        for (int g0 = 0; g0 < nsb; g0 += kNBG) {
            const int nb = (nsb - g0) < kNBG ? (nsb - g0) : kNBG;
            for (int e = tid; e < ne; e += WG) {
                const int se = a.ent_side[sd.ent_off + e];
                const double pxe = a.ent_x[sd.ent_off + e], pye = a.ent_y[sd.ent_off + e];
                const double cs = sd.cshape[se];
                for (int b = 0; b < nb; ++b) {
                    const int p0 = (g0 + b) * S;
                    const int p1 = (p0 + S) < it.n ? (p0 + S) : it.n;
                    double acc = 0.0;
                    for (int p = p0; p < p1; ++p)
                        if (rside[p] == se) acc = acc + rY[p] * rbf(pxe, pye, rx[p], ry[p], cs);
                    Phi[b * a.ne_cap + e] = acc;
                }
            }
            for (int b = tid; b < nb; b += WG) {
                const int p0 = (g0 + b) * S;
                const int p1 = (p0 + S) < it.n ? (p0 + S) : it.n;
                double bs = 0.0, bb = 0.0;
                for (int p = p0; p < p1; ++p) { double v = rs[p]; bs = bs + v; bb = bb + v * v; }
                s_bsb[b] = bs;
                s_bbsb[b] = bb;
            }
            __syncthreads();
            for (int c = tid; c < nnz; c += WG) {
                double G[kNBG];
#pragma unroll
                for (int b = 0; b < kNBG; ++b) G[b] = 0.0;
                const int ec[2] = {a.csr_e1[sd.csr_off + c], a.csr_e2[sd.csr_off + c]};
                for (int q = 0; q < 2; ++q) {
                    const int e = ec[q];
                    if (e < 0) continue;
                    const int se = a.ent_side[sd.ent_off + e];
                    const int m = sd.msz[se];
                    const int el = e - eoff[se];
                    const double *A = a.ipsi + sd.ipsi_off[se];
                    double H[kNBG];
#pragma unroll
                    for (int b = 0; b < kNBG; ++b) H[b] = 0.0;
                    for (int j = 0; j < m; ++j) {
                        const double aj = A[(int64_t)j * m + el];
                        const double *ph = Phi + eoff[se] + j;
#pragma unroll
                        for (int b = 0; b < kNBG; ++b)
                            if (b < nb) H[b] = H[b] + aj * ph[b * a.ne_cap];
                    }
#pragma unroll
                    for (int b = 0; b < kNBG; ++b) G[b] = G[b] + (-H[b]);
                }
                double sum = accS[c], q2 = accQ[c];
                for (int b = 0; b < nb; ++b) {
                    const int p0 = (g0 + b) * S;
                    const int p1 = (p0 + S) < it.n ? (p0 + S) : it.n;
                    sum = sum + G[b];
                    q2 = q2 + (G[b] * G[b]) / (double)(p1 - p0);
                }
                accS[c] = sum;
                accQ[c] = q2;
            }
            if (tid == 0)
                for (int b = 0; b < nb; ++b) { Bs = Bs + s_bsb[b]; BBs = BBs + s_bbsb[b]; }
            __syncthreads();
        }
        for (int c = tid; c < nnz; c += WG) {
            a.partials[it.part_off + c] = accS[c];
            a.partials[it.part_off + nnz + c] = accQ[c];
        }
        if (tid == 0) {
            ItemStat st;
            st.Bs = Bs;
            st.BBs = BBs;
            st.steps = s_steps;
            for (int q = 0; q < 6; ++q) st.cnt[q] = s_cnt[q];
            st.capped = s_cnt[6];
            st.pad = 0;
            a.stats[item] = st;
        }
        __syncthreads();
        if (!PERSIST) break;
    }
}

// One block per knot: sum the knot's items in item order (defined in engine_f64.hip).
__global__ void finalize_kernel(FinalizeArgs f);

// ----------------------------------------------------------------------------
// Per-path kernel (MCinCUDA shape): one path per lane, records written out.
// ----------------------------------------------------------------------------
template <class R, int PDE>
__global__ void __launch_bounds__(256) paths_kernel(PathsArgs a) {
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
    pde_coef<PDE, R>(s.x, s.y, a.pp, co);
    R ex = s.x, ey = s.y;
    for (int phase = 0; phase < (a.phase2 ? 2 : 1); ++phase) {
        const R beta = phase == 0 ? (R)a.beta : (R)0;
        int64_t st0 = s.st;
        bool go = true;
        while (go) {
            R n1, n2, nx, ny;
            int kind;
            draw_normals<R>(s, a.seed, a.stream, a.knot, n1, n2);
            em_step<R>(s, co, h, sqrth, n1, n2);
            s.st += 1;
            R d = rect_dist<R>(rect, s.x, s.y, ex, ey, nx, ny, kind);
            pde_coef<PDE, R>(s.x, s.y, a.pp, co);
            R ga = nx * co.s00 + ny * co.s10;
            R gb = nx * co.s01 + ny * co.s11;
            R nrm = RM<R>::sqrt(ga * ga + gb * gb);
            go = (d < beta * nrm * sqrth) && (s.st - st0 < a.max_steps);
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
        if (os) os[i] = s.st - st0;
    }
}

// Device LUT evaluation for tests (FP64 arithmetic, field f)
__global__ void lut_eval_kernel(LutView L, int field, int64_t n, const double *x, const double *y, double *out);

}  // namespace pdds
