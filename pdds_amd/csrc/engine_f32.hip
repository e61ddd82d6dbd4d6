// engine_f32.hip -- FP32 path-state instantiations: X, Y, Z, xi in FP32 with
// Kahan-compensated Z and xi, hardware transcendentals, one Philox draw per
// two steps; exit records, moments and CSR partials stay in FP64.
#include "engine_kernels.h"

namespace pdds {

template <int PDE, int GEOM, int FAM>
static hipError_t path_f32_gf(bool persistent, bool mixed, int grid, int wg, hipStream_t s, const PathKernelArgs &a) {
    if (!persistent)
        hipLaunchKernelGGL((plain_kernel<float, PDE, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    else if (mixed)
        hipLaunchKernelGGL((path_kernel<float, PDE, true, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    else
        hipLaunchKernelGGL((path_kernel<float, PDE, false, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

// the default control-variate step has kernels of its own, without a family test in the loop
template <int PDE, int GEOM>
static hipError_t path_f32_g(bool persistent, bool mixed, int grid, int wg, hipStream_t s, const PathKernelArgs &a) {
    return a.pp.step == STEP_CV ? path_f32_gf<PDE, GEOM, STEP_CV>(persistent, mixed, grid, wg, s, a)
                                : path_f32_gf<PDE, GEOM, STEP_ANY>(persistent, mixed, grid, wg, s, a);
}

template <int PDE>
static hipError_t path_f32(int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                          const PathKernelArgs &a) {
    if (a.pp.refl) {
        // reflecting sides: rectangles, one stream and k2 per launch, every step family
        if (mixed || geom != GeomRect) return hipErrorInvalidValue;
        if (!persistent)
            hipLaunchKernelGGL((plain_kernel<float, PDE, GeomRect, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
        else
            hipLaunchKernelGGL((path_kernel<float, PDE, false, GeomRect, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
        return hipGetLastError();
    }
    return geom == GeomDisc ? path_f32_g<PDE, GeomDisc>(persistent, mixed, grid, wg, s, a)
                            : path_f32_g<PDE, GeomRect>(persistent, mixed, grid, wg, s, a);
}

hipError_t launch_path_f32(int pde, int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                           const PathKernelArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return path_f32<PDE_MONEGROS>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_SINSIN: return path_f32<PDE_SINSIN>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF: return path_f32<PDE_VARCOEF>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF_LUT: return path_f32<PDE_VARCOEF_LUT>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_HELMHOLTZ: return path_f32<PDE_HELMHOLTZ>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_TABLE: return path_f32<PDE_TABLE>(geom, persistent, mixed, grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

template <bool MIXED, int GEOM, int FAM>
static hipError_t path_tiled_f32_f(int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)path_kernel_tiled<float, MIXED, GEOM, FAM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((path_kernel_tiled<float, MIXED, GEOM, FAM>), dim3(grid), dim3(tiled_wg<float, FAM, GEOM>()), lds, s, a);
    return hipGetLastError();
}

template <bool MIXED, int GEOM>
static hipError_t path_tiled_f32(int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    return a.pp.step == STEP_CV ? path_tiled_f32_f<MIXED, GEOM, STEP_CV>(grid, lds, s, a)
                                : path_tiled_f32_f<MIXED, GEOM, STEP_ANY>(grid, lds, s, a);
}

hipError_t launch_path_tiled_f32(int geom, bool mixed, int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    if (geom == GeomDisc)
        return mixed ? path_tiled_f32<true, GeomDisc>(grid, lds, s, a) : path_tiled_f32<false, GeomDisc>(grid, lds, s, a);
    return mixed ? path_tiled_f32<true, GeomRect>(grid, lds, s, a) : path_tiled_f32<false, GeomRect>(grid, lds, s, a);
}

template <int PDE>
static hipError_t paths_f32(int grid, int wg, hipStream_t s, const PathsArgs &a) {
    if (a.pp.refl)
        hipLaunchKernelGGL((paths_kernel<float, PDE, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
    else if (a.pp.step == STEP_CV)
        hipLaunchKernelGGL((paths_kernel<float, PDE, STEP_CV>), dim3(grid), dim3(wg), 0, s, a);
    else
        hipLaunchKernelGGL((paths_kernel<float, PDE, STEP_ANY>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_paths_f32(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return paths_f32<PDE_MONEGROS>(grid, wg, s, a);
        case PDE_SINSIN: return paths_f32<PDE_SINSIN>(grid, wg, s, a);
        case PDE_VARCOEF: return paths_f32<PDE_VARCOEF>(grid, wg, s, a);
        case PDE_VARCOEF_LUT: return paths_f32<PDE_VARCOEF_LUT>(grid, wg, s, a);
        case PDE_HELMHOLTZ: return paths_f32<PDE_HELMHOLTZ>(grid, wg, s, a);
        case PDE_TABLE: return paths_f32<PDE_TABLE>(grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pdds
