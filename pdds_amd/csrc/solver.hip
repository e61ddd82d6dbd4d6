// solver.hip -- device BiCGSTAB for (I + G) u = B (PDDAlgorithms/PDDSparseGM.cpp:3168-3246 solves the
// same system with Eigen's BiCGSTAB, whose Jacobi preconditioner is the identity here because diag(I+G) = 1)
// and the host LUT prefilter.
#include <cmath>
#include <vector>

#include "../../include/pdds_engine.h"
#include "solver.h"

namespace pdds {

constexpr int kSolveWG = 1024;

// y = (I + G) x, rows strided over the work-group
__device__ void spmv_ig(int n, const int64_t *rp, const int32_t *ci, const double *v, const double *x, double *y) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = x[i];
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s = s + v[k] * x[ci[k]];
        y[i] = s;
    }
}

// fixed-order block dot product (per-thread strided partials, then a tree in LDS)
__device__ double block_dot(int n, const double *a, const double *b, double *red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s = s + a[i] * b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kSolveWG)
bicgstab_kernel(int n, const int64_t *rp, const int32_t *ci, const double *v, const double *b, double *x, double *work,
                double tol, int maxit, int *iters, double *relres) {
    __shared__ double red[kSolveWG];
    double *r = work, *r0 = work + n, *p = work + 2 * (size_t)n, *vv = work + 3 * (size_t)n, *s = work + 4 * (size_t)n,
           *t = work + 5 * (size_t)n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        x[i] = 0.0; r[i] = b[i]; r0[i] = b[i]; p[i] = 0.0; vv[i] = 0.0;
    }
    __syncthreads();
    const double bn2 = block_dot(n, b, b, red);
    if (bn2 == 0.0) {
        if (threadIdx.x == 0) { *iters = 0; *relres = 0.0; }
        return;
    }
    double rho = 1.0, alpha = 1.0, w = 1.0, r0n2 = bn2;
    const double tol2 = tol * tol * bn2;
    int it = 0;
    double rn2 = block_dot(n, r, r, red);
    while (rn2 > tol2 && it < maxit) {
        const double rho_old = rho;
        rho = block_dot(n, r0, r, red);
        if (fabs(rho) < 1e-30 * r0n2) {
            // breakdown: restart from the current residual
            spmv_ig(n, rp, ci, v, x, t);
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += blockDim.x) { r[i] = b[i] - t[i]; r0[i] = r[i]; }
            __syncthreads();
            rho = r0n2 = block_dot(n, r, r, red);
        }
        const double beta = (rho / rho_old) * (alpha / w);
        for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = r[i] + beta * (p[i] - w * vv[i]);
        __syncthreads();
        spmv_ig(n, rp, ci, v, p, vv);
        __syncthreads();
        alpha = rho / block_dot(n, r0, vv, red);
        for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = r[i] - alpha * vv[i];
        __syncthreads();
        spmv_ig(n, rp, ci, v, s, t);
        __syncthreads();
        const double tt = block_dot(n, t, t, red);
        w = tt > 0.0 ? block_dot(n, t, s, red) / tt : 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            x[i] = x[i] + (alpha * p[i] + w * s[i]);
            r[i] = s[i] - w * t[i];
        }
        __syncthreads();
        rn2 = block_dot(n, r, r, red);
        ++it;
    }
    if (threadIdx.x == 0) { *iters = it; *relres = sqrt(rn2 / bn2); }
}

hipError_t bicgstab_device(int n, const int64_t *row_ptr, const int32_t *cols, const double *vals, const double *b,
                           double *x, double *work, double tol, int maxit, int *iters_dev, double *relres_dev,
                           hipStream_t s) {
    hipLaunchKernelGGL(bicgstab_kernel, dim3(1), dim3(kSolveWG), 0, s, n, row_ptr, cols, vals, b, x, work, tol, maxit,
                       iters_dev, relres_dev);
    return hipGetLastError();
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

