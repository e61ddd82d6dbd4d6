// engine_f64.hip -- FP64 instantiations of the path kernels, and the reduction,
// finalize and LUT-evaluation kernels (shared by both path precisions).  Built
// with -ffp-contract=off so that the only fused multiply-adds are the explicit
// fma() calls: every FP64 path and every reduction is then bit-identical to the
// FP64 CPU restatement on the same Philox stream.
#include "../../include/pdds_engine.h"
#include "engine_kernels.h"

namespace pdds {

template <int PDE>
static hipError_t path_f64(bool persistent, bool mixed, int grid, int wg, hipStream_t s, const PathKernelArgs &a) {
    if (!persistent)
        hipLaunchKernelGGL((plain_kernel<double, PDE>), dim3(grid), dim3(wg), 0, s, a);
    else if (mixed)
        hipLaunchKernelGGL((path_kernel<double, PDE, true>), dim3(grid), dim3(wg), 0, s, a);
    else
        hipLaunchKernelGGL((path_kernel<double, PDE, false>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_path_f64(int pde, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                           const PathKernelArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return path_f64<PDE_MONEGROS>(persistent, mixed, grid, wg, s, a);
        case PDE_SINSIN: return path_f64<PDE_SINSIN>(persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF: return path_f64<PDE_VARCOEF>(persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF_LUT: return path_f64<PDE_VARCOEF_LUT>(persistent, mixed, grid, wg, s, a);
        case PDE_HELMHOLTZ: return path_f64<PDE_HELMHOLTZ>(persistent, mixed, grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

int path_tiled_wg(int precision) { return precision == PDDS_FP64 ? tiled_wg<double>() : tiled_wg<float>(); }

template <bool MIXED>
static hipError_t path_tiled_f64(int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)path_kernel_tiled<double, MIXED>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((path_kernel_tiled<double, MIXED>), dim3(grid), dim3(tiled_wg<double>()), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_path_tiled_f64(bool mixed, int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    return mixed ? path_tiled_f64<true>(grid, lds, s, a) : path_tiled_f64<false>(grid, lds, s, a);
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

// One work-group per work item: per-side RBF moments per sub-batch, G_beta = -Psi^-T Phi per
// CSR column, sums in fixed order (sub-batch -> item), B sums per sub-batch.  A group of kNBG
// sub-batches is staged in LDS with, per (sub-batch, side), the list of its paths that exit
// through that side in path order; the moment of entry e in sub-batch b then loops over that
// list only, adding in path order (the order of the per-exit loop it replaces).
__global__ void __launch_bounds__(256) reduce_kernel(ReduceArgs a) {
    extern __shared__ double smem[];
    __shared__ double s_bsb[kNBG], s_bbsb[kNBG];
    __shared__ int s_lcnt[kNBG * 4];
    __shared__ unsigned long long s_steps;
    __shared__ unsigned int s_cnt[8];
    const int tid = threadIdx.x;
    const int WG = blockDim.x;
    const int64_t item = a.item_base + (int64_t)blockIdx.x;
    const ItemDesc it = a.items[item];
    const KnotDesc kd = a.knots[it.knot];
    const TableDesc tb = a.tables[kd.table];
    const StencilDesc &sd = tb.sten[kd.stencil];
    const int64_t g0 = kd.path_off + it.p_begin;
    const double *rx = a.rec.x + g0, *ry = a.rec.y + g0, *rY = a.rec.Y + g0, *rs = a.rec.score + g0;
    const int8_t *rside = a.rec.side + g0;
    const int32_t *rst = a.rec.steps + g0;
    const int S = a.S;
    const int GS = kNBG * S;
    double *Phi = smem;
    double *accS = smem + kNBG * a.ne_cap;
    double *accQ = accS + a.nnz_cap;
    double *gx = accQ + a.nnz_cap;
    double *gy = gx + GS;
    double *gY = gy + GS;
    uint16_t *lst = (uint16_t *)(gY + GS);
    int8_t *gside = (int8_t *)(lst + 4 * GS);
    if (tid == 0) {
        s_steps = 0ULL;
        for (int q = 0; q < 8; ++q) s_cnt[q] = 0u;
    }
    __syncthreads();
    {
        unsigned long long my = 0ULL;
        unsigned int cnt[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int p = tid; p < it.n; p += WG) {
            const int32_t st = rst[p];
            my += (unsigned long long)(st & ~kCappedBit);
            if (st & kCappedBit) cnt[6]++;
            const int sd0 = rside[p];
            cnt[sd0 < 0 ? 5 : sd0]++;
        }
        atomicAdd(&s_steps, my);
        for (int q = 0; q < 7; ++q)
            if (cnt[q]) atomicAdd(&s_cnt[q], cnt[q]);
    }
    const int ne = sd.msz[0] + sd.msz[1] + sd.msz[2] + sd.msz[3];
    const int nnz = sd.csr_nnz;
    const int nsb = (it.n + S - 1) / S;
    int eoff[5];
    eoff[0] = 0;
    for (int q = 0; q < 4; ++q) eoff[q + 1] = eoff[q] + sd.msz[q];
    for (int c = tid; c < nnz; c += WG) { accS[c] = 0.0; accQ[c] = 0.0; }
    double Bs = 0.0, BBs = 0.0;
    for (int g = 0; g < nsb; g += kNBG) {
        const int nb = (nsb - g) < kNBG ? (nsb - g) : kNBG;
        const int q0 = g * S;                                   // first path of the group
        const int gn = (it.n - q0) < nb * S ? (it.n - q0) : nb * S;
        for (int p = tid; p < gn; p += WG) {
            gx[p] = rx[q0 + p];
            gy[p] = ry[q0 + p];
            gY[p] = rY[q0 + p];
            gside[p] = rside[q0 + p];
        }
        __syncthreads();
        // per (sub-batch, side) lists in path order, and the B sums of each sub-batch
        if (tid < nb * 4) {
            const int b = tid >> 2, q = tid & 3;
            const int p0 = b * S, p1 = (p0 + S) < gn ? (p0 + S) : gn;
            int c = 0;
            for (int p = p0; p < p1; ++p)
                if (gside[p] == q) lst[(b * 4 + q) * S + c++] = (uint16_t)(p - p0);
            s_lcnt[b * 4 + q] = c;
        } else if (tid >= 64 && tid < 64 + nb) {
            const int b = tid - 64;
            const int p0 = q0 + b * S;
            const int p1 = (p0 + S) < it.n ? (p0 + S) : it.n;
            double bs = 0.0, bb = 0.0;
            for (int p = p0; p < p1; ++p) { const double v = rs[p]; bs = bs + v; bb = bb + v * v; }
            s_bsb[b] = bs;
            s_bbsb[b] = bb;
        }
        __syncthreads();
        // moments: unit u = (sub-batch b, entry e), entries fastest so that a wave reads one list
        for (int u = tid; u < nb * ne; u += WG) {
            const int b = u / ne, e = u - b * ne;
            const int se = tb.ents[sd.ent_off + e];
            const double pxe = tb.entx[sd.ent_off + e], pye = tb.enty[sd.ent_off + e];
            const double ic2 = sd.inv_c2[se];
            const uint16_t *L = lst + (b * 4 + se) * S;
            const int nl = s_lcnt[b * 4 + se];
            const int pb = b * S;
            double acc = 0.0;
            for (int l = 0; l < nl; ++l) {
                const int p = pb + L[l];
                acc = acc + gY[p] * rbf(pxe, pye, gx[p], gy[p], ic2);
            }
            Phi[b * a.ne_cap + e] = acc;
        }
        __syncthreads();
        for (int c = tid; c < nnz; c += WG) {
            double G[kNBG];
#pragma unroll
            for (int b = 0; b < kNBG; ++b) G[b] = 0.0;
            const int ec[2] = {tb.e1[sd.csr_off + c], tb.e2[sd.csr_off + c]};
            for (int q = 0; q < 2; ++q) {
                const int e = ec[q];
                if (e < 0) continue;
                const int se = tb.ents[sd.ent_off + e];
                const int m = sd.msz[se];
                const int el = e - eoff[se];
                const double *A = tb.ipsi + sd.ipsi_off[se];
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
                const int p0 = (g + b) * S;
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
}

hipError_t launch_reduce(int grid, int wg, size_t shmem, hipStream_t s, const ReduceArgs &a) {
    hipLaunchKernelGGL(reduce_kernel, dim3(grid), dim3(wg), shmem, s, a);
    return hipGetLastError();
}

// One block per knot: sum the knot's items in item order.
__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f) {
    const int k = blockIdx.x;
    if (k >= f.n_knots) return;
    const KnotDesc kd = f.knots[k];
    const StencilDesc &sd = f.tables[kd.table].sten[kd.stencil];
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
    if (L.nf > 5) lut_eval<double, 8>(L, x[i], y[i], v);
    else if (L.nf == 5) lut_eval<double, 5>(L, x[i], y[i], v);
    else lut_eval<double, 4>(L, x[i], y[i], v);
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
