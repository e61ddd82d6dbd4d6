// solver.hip -- device BiCGSTAB for (I + G) u = B (PDDAlgorithms/PDDSparseGM.cpp:3168-3246 solves the
// same system with Eigen's BiCGSTAB, whose Jacobi preconditioner is the identity here because diag(I+G) = 1)
// and the host LUT prefilter.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/pdds_engine.h"
#include "solver.h"

namespace pdds {

// One wave per row for the products with I + G (rows of G carry ~10^2..10^3 entries), four rows
// per work-group; elementwise updates 256 per work-group.  Every reduction has a fixed order
// (wave xor-tree, then work-group partials summed in index order by every consumer), so the
// iterates do not depend on scheduling.  Iteration k reads its scalars from slot k-1 and writes
// slot k, so no kernel reads a scalar another work-group of the same launch may be writing.
constexpr int kRowsPerWG = 4;
constexpr int kEwWG = 256;

enum { ST_RUN = 0, ST_DONE = 1, ST_BREAKDOWN = 2 };

struct SolveSlot {
    double rho, alpha, omega;   // BiCGSTAB scalars after this iteration
    double rr;                  // ||r||^2 entering this iteration
    double relres;              // final relative residual (when done)
    int32_t status, iters;      // ST_*, iterations done (when done or broken down)
};

struct SolveGlobal { double bn2, tol2, r0n2; };

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
    return v;
}

// sum of p[0..np) in a fixed order; every thread of the work-group gets the result
__device__ double group_sum(const double *p, int np, double *red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s = s + p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// (G y)_i for the row of this wave, y given elementwise by a functor
template <class F>
__device__ __forceinline__ double row_gy(const int64_t *rp, const int32_t *ci, const double *v, int i, int lane,
                                         F y) {
    double s = 0.0;
    for (int64_t k = rp[i] + lane; k < rp[i + 1]; k += 64) s = s + v[k] * y(ci[k]);
    return wave_sum(s);
}

struct SolveArgs {
    int n;
    const int64_t *rp;
    const int32_t *ci;
    const double *v, *b;
    double *x, *r, *rh, *P[2], *V[2], *S, *T;
    double *partA, *partB, *partC;   // [nbA], [2][nbA], [2][nbC]
    int nbA, nbC;
    SolveSlot *slot;
    SolveGlobal *glob;
    int maxit;
};

// r = b - (I + G) x, rh = r, p = v = 0; partials of r.r (and b.b on the first start) into partB
__global__ void __launch_bounds__(64 * kRowsPerWG) solve_init(SolveArgs a, int first) {
    __shared__ double sr[kRowsPerWG], sb[kRowsPerWG];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kRowsPerWG + w;
    double rr = 0.0, bb = 0.0;
    if (i < a.n) {
        const double *x = a.x;
        const double gx = row_gy(a.rp, a.ci, a.v, i, lane, [&](int j) { return x[j]; });
        const double ri = a.b[i] - (x[i] + gx);
        if (lane == 0) {
            a.r[i] = ri; a.rh[i] = ri;
            a.P[0][i] = 0.0; a.P[1][i] = 0.0; a.V[0][i] = 0.0; a.V[1][i] = 0.0;
        }
        rr = ri * ri;
        bb = a.b[i] * a.b[i];
    }
    if (lane == 0) { sr[w] = rr; sb[w] = bb; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s1 = 0.0, s2 = 0.0;
        for (int q = 0; q < kRowsPerWG; ++q) { s1 = s1 + sr[q]; s2 = s2 + sb[q]; }
        a.partB[blockIdx.x] = s1;
        if (first) a.partB[a.nbA + blockIdx.x] = s2;
    }
}

// slot k0 = a fresh start: rho = alpha = omega = 1; ||b||^2 and the tolerance on the first start
__global__ void __launch_bounds__(256) solve_init_finish(SolveArgs a, int k0, double tol, int first) {
    __shared__ double red[256];
    const double rr = group_sum(a.partB, a.nbA, red);
    double bn2 = 0.0;
    if (first) bn2 = group_sum(a.partB + a.nbA, a.nbA, red);
    if (threadIdx.x == 0) {
        if (first) { a.glob->bn2 = bn2; a.glob->tol2 = tol * tol * bn2; }
        a.glob->r0n2 = rr;
        SolveSlot s;
        s.rho = 1.0; s.alpha = 1.0; s.omega = 1.0; s.rr = rr; s.relres = 0.0;
        s.status = ST_RUN; s.iters = k0;
        a.slot[k0] = s;
        // the first K_A after a start reduces r.rh = r.r from partC: copy the partials there
    }
    for (int q = threadIdx.x; q < a.nbA; q += blockDim.x) { a.partC[q] = a.partB[q]; a.partC[a.nbA + q] = a.partB[q]; }
}

// iteration k, part A: rho = rh.r, convergence test, p = r + beta (p - omega v), v = (I+G) p
__global__ void __launch_bounds__(64 * kRowsPerWG) solve_ka(SolveArgs a, int k, int npc) {
    __shared__ double red[64 * kRowsPerWG];
    __shared__ double sv[kRowsPerWG];
    const SolveSlot prev = a.slot[k - 1];
    if (prev.status != ST_RUN) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.slot[k] = prev;
        return;
    }
    // partC holds rh.r in [0, npc) and r.r in [npc, 2 npc) of the previous update
    const double rho = group_sum(a.partC, npc, red);
    const double rr = group_sum(a.partC + npc, npc, red);
    const SolveGlobal gl = *a.glob;
    const bool conv = !(rr > gl.tol2) || gl.bn2 == 0.0;
    const bool maxed = k - 1 >= a.maxit;
    const bool brk = !conv && !maxed && fabs(rho) < 1e-30 * gl.r0n2;
    if (conv || maxed || brk) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            SolveSlot s = prev;
            s.rr = rr;
            s.status = brk ? ST_BREAKDOWN : ST_DONE;
            s.iters = k - 1;
            s.relres = gl.bn2 > 0.0 ? sqrt(rr / gl.bn2) : 0.0;
            a.slot[k] = s;
        }
        return;
    }
    const double beta = (rho / prev.rho) * (prev.alpha / prev.omega);
    const double om = prev.omega;
    const double *r = a.r, *po = a.P[(k - 1) & 1], *vo = a.V[(k - 1) & 1];
    auto pnew = [&](int j) { return r[j] + beta * (po[j] - om * vo[j]); };
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kRowsPerWG + w;
    double c = 0.0;
    if (i < a.n) {
        const double pi = pnew(i);
        const double vi = pi + row_gy(a.rp, a.ci, a.v, i, lane, pnew);
        if (lane == 0) { a.P[k & 1][i] = pi; a.V[k & 1][i] = vi; }
        c = a.rh[i] * vi;
    }
    if (lane == 0) sv[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < kRowsPerWG; ++q) s = s + sv[q];
        a.partA[blockIdx.x] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SolveSlot s = prev;
        s.rho = rho; s.rr = rr; s.status = ST_RUN;
        a.slot[k] = s;   // alpha, omega filled in by parts B and C
    }
}

// part B: alpha = rho / rh.v, s = r - alpha v, t = (I+G) s
__global__ void __launch_bounds__(64 * kRowsPerWG) solve_kb(SolveArgs a, int k) {
    __shared__ double red[64 * kRowsPerWG];
    __shared__ double s1[kRowsPerWG], s2[kRowsPerWG];
    // only the fields earlier kernels wrote: block 0 of this launch writes slot k's alpha
    if (a.slot[k].status != ST_RUN) return;
    const double rv = group_sum(a.partA, a.nbA, red);
    const double alpha = a.slot[k].rho / rv;
    const double *r = a.r, *vn = a.V[k & 1];
    auto sval = [&](int j) { return r[j] - alpha * vn[j]; };
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kRowsPerWG + w;
    double ts = 0.0, tt = 0.0;
    if (i < a.n) {
        const double si = sval(i);
        const double ti = si + row_gy(a.rp, a.ci, a.v, i, lane, sval);
        if (lane == 0) { a.S[i] = si; a.T[i] = ti; }
        ts = ti * si;
        tt = ti * ti;
    }
    if (lane == 0) { s1[w] = ts; s2[w] = tt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double u1 = 0.0, u2 = 0.0;
        for (int q = 0; q < kRowsPerWG; ++q) { u1 = u1 + s1[q]; u2 = u2 + s2[q]; }
        a.partB[blockIdx.x] = u1;
        a.partB[a.nbA + blockIdx.x] = u2;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.slot[k].alpha = alpha;
}

// part C: omega = t.s / t.t, x += alpha p + omega s, r = s - omega t; partials of rh.r and r.r
__global__ void __launch_bounds__(kEwWG) solve_kc(SolveArgs a, int k) {
    __shared__ double red[kEwWG];
    // only the fields earlier kernels wrote: block 0 of this launch writes slot k's omega
    if (a.slot[k].status != ST_RUN) return;
    const double ts = group_sum(a.partB, a.nbA, red);
    const double tt = group_sum(a.partB + a.nbA, a.nbA, red);
    const double om = tt > 0.0 ? ts / tt : 0.0;
    const double alpha = a.slot[k].alpha;
    const int i = blockIdx.x * kEwWG + threadIdx.x;
    double c1 = 0.0, c2 = 0.0;
    if (i < a.n) {
        const double si = a.S[i];
        a.x[i] = a.x[i] + (alpha * a.P[k & 1][i] + om * si);
        const double ri = si - om * a.T[i];
        a.r[i] = ri;
        c1 = a.rh[i] * ri;
        c2 = ri * ri;
    }
    red[threadIdx.x] = c1;
    __syncthreads();
    for (int w = kEwWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partC[blockIdx.x] = red[0];
    __syncthreads();
    red[threadIdx.x] = c2;
    __syncthreads();
    for (int w = kEwWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partC[a.nbC + blockIdx.x] = red[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.slot[k].omega = om;
}

__global__ void solve_zero(double *x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

size_t bicgstab_work_bytes(int n, int maxit) {
    const size_t nbA = ((size_t)n + kRowsPerWG - 1) / kRowsPerWG, nbC = ((size_t)n + kEwWG - 1) / kEwWG;
    const size_t nbP = nbA > nbC ? nbA : nbC;
    return (size_t)8 * n * sizeof(double) + (nbA + 2 * nbA + 2 * nbP) * sizeof(double) +
           ((size_t)maxit + 2) * sizeof(SolveSlot) + sizeof(SolveGlobal) + 64;
}

hipError_t bicgstab_device(int n, const int64_t *row_ptr, const int32_t *cols, const double *vals, const double *b,
                           double *x, void *work, double tol, int maxit, int batch, int *iters, double *relres,
                           hipStream_t st) {
    SolveArgs a;
    a.n = n; a.rp = row_ptr; a.ci = cols; a.v = vals; a.b = b; a.x = x;
    a.nbA = (n + kRowsPerWG - 1) / kRowsPerWG;
    a.nbC = (n + kEwWG - 1) / kEwWG;
    const int nbP = a.nbA > a.nbC ? a.nbA : a.nbC;
    double *w = (double *)work;
    a.r = w; a.rh = w + n; a.P[0] = w + 2 * (size_t)n; a.P[1] = w + 3 * (size_t)n;
    a.V[0] = w + 4 * (size_t)n; a.V[1] = w + 5 * (size_t)n; a.S = w + 6 * (size_t)n; a.T = w + 7 * (size_t)n;
    a.partA = w + 8 * (size_t)n;
    a.partB = a.partA + a.nbA;
    a.partC = a.partB + 2 * (size_t)a.nbA;
    a.slot = (SolveSlot *)(a.partC + 2 * (size_t)nbP);
    a.glob = (SolveGlobal *)(a.slot + (maxit + 2));
    a.maxit = maxit;
    if (batch < 1) batch = 8;
    hipLaunchKernelGGL(solve_zero, dim3((n + 255) / 256), dim3(256), 0, st, x, n);
    int k0 = 0;          // slot of the current (re)start
    bool first = true;
    while (true) {
        hipLaunchKernelGGL(solve_init, dim3(a.nbA), dim3(64 * kRowsPerWG), 0, st, a, first ? 1 : 0);
        hipLaunchKernelGGL(solve_init_finish, dim3(1), dim3(256), 0, st, a, k0, tol, first ? 1 : 0);
        first = false;
        int k = k0 + 1, npc = a.nbA;   // after a start, partC holds nbA partials
        SolveSlot last;
        while (true) {
            const int kend = std::min(k + batch, maxit + 2);
            for (; k < kend; ++k) {
                hipLaunchKernelGGL(solve_ka, dim3(a.nbA), dim3(64 * kRowsPerWG), 0, st, a, k, npc);
                hipLaunchKernelGGL(solve_kb, dim3(a.nbA), dim3(64 * kRowsPerWG), 0, st, a, k);
                hipLaunchKernelGGL(solve_kc, dim3(a.nbC), dim3(kEwWG), 0, st, a, k);
                npc = a.nbC;
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            e = hipMemcpyAsync(&last, a.slot + (k - 1), sizeof(SolveSlot), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            if (last.status != ST_RUN || k >= maxit + 2) break;
        }
        if (last.status == ST_BREAKDOWN && last.iters < maxit) {
            k0 = last.iters;   // restart from the current x at the iteration that broke down
            continue;
        }
        *iters = last.iters;
        *relres = last.relres;
        return hipSuccess;
    }
}

static void prefilter_1d(int n, const double *s, double *c) {
    // end conditions: spline second difference at the end node = extrapolated data second difference
    const double rL = 2.0 * (s[0] - 2.0 * s[1] + s[2]) - (s[1] - 2.0 * s[2] + s[3]);
    const double rR = 2.0 * (s[n - 1] - 2.0 * s[n - 2] + s[n - 3]) - (s[n - 2] - 2.0 * s[n - 3] + s[n - 4]);
    std::vector<double> cc(n), cp(n), dp(n);
    cc[0] = s[0] - rL / 6.0;
    cc[n - 1] = s[n - 1] - rR / 6.0;
    const int m = n - 2;  // unknowns cc[1..n-2]: cc[k-1] + 4 cc[k] + cc[k+1] = 6 s[k]  (Thomas algorithm)
    for (int k = 0; k < m; ++k) {
        double rhs = 6.0 * s[k + 1];
        if (k == 0) rhs -= cc[0];
        if (k == m - 1) rhs -= cc[n - 1];
        const double den = 4.0 - (k > 0 ? cp[k - 1] : 0.0);
        cp[k] = 1.0 / den;
        dp[k] = (rhs - (k > 0 ? dp[k - 1] : 0.0)) / den;
    }
    for (int k = m - 1; k >= 0; --k) cc[k + 1] = dp[k] - cp[k] * (k + 1 < m ? cc[k + 2] : 0.0);
    c[0] = 2.0 * cc[0] - cc[1] + rL;
    for (int k = 0; k < n; ++k) c[k + 1] = cc[k];
    c[n + 1] = 2.0 * cc[n - 1] - cc[n - 2] + rR;
}

void lut_prefilter(int nx, int ny, int nf, const double *samples, float *coef) {
    const int NX = nx + 2, NY = ny + 2;
    std::vector<double> tmp((size_t)NX * ny), row(nx), colv(ny), colc(NY), rowc(NX);
    for (int f = 0; f < nf; ++f) {
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) row[i] = samples[((size_t)j * nx + i) * nf + f];
            prefilter_1d(nx, row.data(), rowc.data());
            for (int i = 0; i < NX; ++i) tmp[(size_t)j * NX + i] = rowc[i];
        }
        for (int i = 0; i < NX; ++i) {
            for (int j = 0; j < ny; ++j) colv[j] = tmp[(size_t)j * NX + i];
            prefilter_1d(ny, colv.data(), colc.data());
            for (int j = 0; j < NY; ++j) coef[((size_t)j * NX + i) * nf + f] = (float)colc[j];
        }
    }
}

}  // namespace pdds

