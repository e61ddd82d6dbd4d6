// engine_f64.hip -- FP64 instantiations of the path kernels.  Built with
// -ffp-contract=off so that the only fused multiply-adds are the explicit
// fma() calls of pdds_math.h: every FP64 path is then bit-identical to the
// FP64 CPU restatement on the same Philox stream.
#include "engine_kernels.h"

namespace pdds {

template <int PDE>
static hipError_t asm_f64(bool persistent, int grid, int wg, size_t shmem, hipStream_t s, const EngineArgs &a) {
    if (persistent)
        hipLaunchKernelGGL((assemble_kernel<double, PDE, true>), dim3(grid), dim3(wg), shmem, s, a);
    else
        hipLaunchKernelGGL((assemble_kernel<double, PDE, false>), dim3(grid), dim3(wg), shmem, s, a);
    return hipGetLastError();
}

hipError_t launch_assemble_f64(int pde, bool persistent, int grid, int wg, size_t shmem, hipStream_t s,
                               const EngineArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return asm_f64<PDE_MONEGROS>(persistent, grid, wg, shmem, s, a);
        case PDE_SINSIN: return asm_f64<PDE_SINSIN>(persistent, grid, wg, shmem, s, a);
        case PDE_VARCOEF: return asm_f64<PDE_VARCOEF>(persistent, grid, wg, shmem, s, a);
        case PDE_VARCOEF_LUT: return asm_f64<PDE_VARCOEF_LUT>(persistent, grid, wg, shmem, s, a);
        case PDE_HELMHOLTZ: return asm_f64<PDE_HELMHOLTZ>(persistent, grid, wg, shmem, s, a);
        default: return hipErrorInvalidValue;
    }
}

template <int PDE>
static hipError_t paths_f64(int grid, int wg, hipStream_t s, const PathsArgs &a) {
    hipLaunchKernelGGL((paths_kernel<double, PDE>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_paths_f64(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return paths_f64<PDE_MONEGROS>(grid, wg, s, a);
        case PDE_SINSIN: return paths_f64<PDE_SINSIN>(grid, wg, s, a);
        case PDE_VARCOEF: return paths_f64<PDE_VARCOEF>(grid, wg, s, a);
        case PDE_VARCOEF_LUT: return paths_f64<PDE_VARCOEF_LUT>(grid, wg, s, a);
        case PDE_HELMHOLTZ: return paths_f64<PDE_HELMHOLTZ>(grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

// One block per knot: sum the knot's items in item order.
__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f) {
    const int k = blockIdx.x;
    if (k >= f.n_knots) return;
    const KnotDesc kd = f.knots[k];
    const StencilDesc &sd = f.sten[kd.stencil];
    const int nnz = sd.csr_nnz;
    const int64_t w0 = kd.item_begin, w1 = kd.item_begin + kd.item_count;
    const double N = (double)kd.n_paths;
    const int64_t nb = (kd.n_paths + f.S - 1) / f.S;
    for (int c = threadIdx.x; c < nnz; c += blockDim.x) {
        double sum = 0.0, q2 = 0.0;
        for (int64_t w = w0; w < w1; ++w) {
            const int64_t off = f.items[w].part_off;
            sum = sum + f.partials[off + c];
            q2 = q2 + f.partials[off + nnz + c];
        }
        const double G = sum / N;
        f.G[kd.out_off + c] = G;
        f.Gvar[kd.out_off + c] = nb > 1 ? (q2 - N * G * G) / (double)(nb - 1) : 0.0;
    }
    if (threadIdx.x == 0) {
        double Bs = 0.0, BBs = 0.0;
        KnotStat ks;
        ks.steps = 0;
        ks.capped = 0;
        for (int q = 0; q < 6; ++q) ks.cnt[q] = 0;
        for (int64_t w = w0; w < w1; ++w) {
            const ItemStat st = f.stats[w];
            Bs = Bs + st.Bs;
            BBs = BBs + st.BBs;
            ks.steps += st.steps;
            ks.capped += st.capped;
            for (int q = 0; q < 6; ++q) ks.cnt[q] += st.cnt[q];
        }
        ks.B = Bs / N;
        ks.BB = BBs / N;
        f.kstats[k] = ks;
    }
}

hipError_t launch_finalize(int grid, int wg, hipStream_t s, const FinalizeArgs &f) {
    hipLaunchKernelGGL(finalize_kernel, dim3(grid), dim3(wg), 0, s, f);
    return hipGetLastError();
}

__global__ void lut_eval_kernel(LutView L, int field, int64_t n, const double *x, const double *y, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[8];
    switch (L.nf) {
        case 1: lut_eval<double, 1>(L, x[i], y[i], v); break;
        case 2: lut_eval<double, 2>(L, x[i], y[i], v); break;
        case 3: lut_eval<double, 3>(L, x[i], y[i], v); break;
        case 4: lut_eval<double, 4>(L, x[i], y[i], v); break;
        case 5: lut_eval<double, 5>(L, x[i], y[i], v); break;
        case 6: lut_eval<double, 6>(L, x[i], y[i], v); break;
        case 7: lut_eval<double, 7>(L, x[i], y[i], v); break;
        default: lut_eval<double, 8>(L, x[i], y[i], v); break;
    }
    out[i] = v[field];
}

hipError_t launch_lut_eval(const LutView &L, int field, int64_t n, const double *x, const double *y, double *out,
                           hipStream_t s) {
    int wg = 256;
    int64_t grid = (n + wg - 1) / wg;
    hipLaunchKernelGGL(lut_eval_kernel, dim3((unsigned)grid), dim3(wg), 0, s, L, field, n, x, y, out);
    return hipGetLastError();
}

}  // namespace pdds
