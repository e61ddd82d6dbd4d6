// solver.h -- interface solve and LUT construction helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdds {

// Interpolating cubic B-spline coefficients of node-aligned samples [ny][nx][nf] -> [(ny+2)][(nx+2)][nf].
// Per axis: (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k] for every node, closed by matching the spline's
// second difference at each end node to the data's second difference extrapolated to that node.
void lut_prefilter(int nx, int ny, int nf, const double *samples, float *coef);

// BiCGSTAB (van der Vorst 1992) for (I + G) u = B with G in CSR, zero initial guess, run on the
// device by three kernels per iteration with fixed-order reductions; the host polls a status word
// every `batch` iterations.  work must hold bicgstab_work_bytes(n, maxit) bytes.  A breakdown
// (rh.r ~ 0) restarts from the current iterate.
size_t bicgstab_work_bytes(int n, int maxit);
hipError_t bicgstab_device(int n, const int64_t *row_ptr, const int32_t *cols, const double *vals, const double *b,
                           double *x, void *work, double tol, int maxit, int batch, int *iters, double *relres,
                           hipStream_t s);

}  // namespace pdds
