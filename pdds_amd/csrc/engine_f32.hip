// engine_f32.hip -- FP32 path-state instantiations: X, Y, Z, xi in FP32 with
// Kahan-compensated Z and xi, hardware transcendentals, one Philox draw per
// two steps; exit records, moments and CSR partials stay in FP64.
#include "engine_kernels.h"

namespace pdds {

template <int PDE>
static hipError_t asm_f32(bool persistent, int grid, int wg, size_t shmem, hipStream_t s, const EngineArgs &a) {
    if (persistent)
        hipLaunchKernelGGL((assemble_kernel<float, PDE, true>), dim3(grid), dim3(wg), shmem, s, a);
    else
        hipLaunchKernelGGL((assemble_kernel<float, PDE, false>), dim3(grid), dim3(wg), shmem, s, a);
    return hipGetLastError();
}

hipError_t launch_assemble_f32(int pde, bool persistent, int grid, int wg, size_t shmem, hipStream_t s,
                               const EngineArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return asm_f32<PDE_MONEGROS>(persistent, grid, wg, shmem, s, a);
        case PDE_SINSIN: return asm_f32<PDE_SINSIN>(persistent, grid, wg, shmem, s, a);
        case PDE_VARCOEF: return asm_f32<PDE_VARCOEF>(persistent, grid, wg, shmem, s, a);
        case PDE_VARCOEF_LUT: return asm_f32<PDE_VARCOEF_LUT>(persistent, grid, wg, shmem, s, a);
        case PDE_HELMHOLTZ: return asm_f32<PDE_HELMHOLTZ>(persistent, grid, wg, shmem, s, a);
        default: return hipErrorInvalidValue;
    }
}

template <int PDE>
static hipError_t paths_f32(int grid, int wg, hipStream_t s, const PathsArgs &a) {
    hipLaunchKernelGGL((paths_kernel<float, PDE>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_paths_f32(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return paths_f32<PDE_MONEGROS>(grid, wg, s, a);
        case PDE_SINSIN: return paths_f32<PDE_SINSIN>(grid, wg, s, a);
        case PDE_VARCOEF: return paths_f32<PDE_VARCOEF>(grid, wg, s, a);
        case PDE_VARCOEF_LUT: return paths_f32<PDE_VARCOEF_LUT>(grid, wg, s, a);
        case PDE_HELMHOLTZ: return paths_f32<PDE_HELMHOLTZ>(grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pdds
