// solver.h -- interface solve and LUT construction helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdds {

// Interpolating cubic B-spline coefficients of node-aligned samples [ny][nx][nf] -> [(ny+2)][(nx+2)][nf].
// Per axis: (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k] for every node, closed by matching the spline's
// second difference at each end node to the data's second difference extrapolated to that node.
void lut_prefilter(int nx, int ny, int nf, const double *samples, float *coef);

// BiCGSTAB (van der Vorst 1992) for (I + G) u = B with G in CSR, zero initial guess, run by one
// resident work-group so the whole iteration stays on the device with fixed-order reductions.
hipError_t bicgstab_device(int n, const int64_t *row_ptr, const int32_t *cols, const double *vals, const double *b,
                           double *x, double *work, double tol, int maxit, int *iters_dev, double *relres_dev,
                           hipStream_t s);

}  // namespace pdds
