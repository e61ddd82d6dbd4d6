// engine_f64.hip -- FP64 instantiations of the path kernels, and the reduction,
// finalize and LUT-evaluation kernels (shared by both path precisions).  Built
// with -ffp-contract=off so that the only fused multiply-adds are the explicit
// fma() calls: every FP64 path and every reduction is then bit-identical to the
// FP64 CPU restatement on the same Philox stream.
#include "../../include/pdds_engine.h"
#include "engine_kernels.h"

namespace pdds {

__constant__ double kExp2TabDev[64] = PDDS_EXP2_TAB;   // 2^(j/64), j = 0..63

template <int PDE, int GEOM, int FAM>
static hipError_t path_f64_gf(bool persistent, bool mixed, int grid, int wg, hipStream_t s, const PathKernelArgs &a) {
    if (!persistent)
        hipLaunchKernelGGL((plain_kernel<double, PDE, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    else if (mixed)
        hipLaunchKernelGGL((path_kernel<double, PDE, true, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    else
        hipLaunchKernelGGL((path_kernel<double, PDE, false, GEOM, FAM>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

// the default control-variate step has kernels of its own, without a family test in the loop
template <int PDE, int GEOM>
static hipError_t path_f64_g(bool persistent, bool mixed, int grid, int wg, hipStream_t s, const PathKernelArgs &a) {
    return a.pp.step == STEP_CV ? path_f64_gf<PDE, GEOM, STEP_CV>(persistent, mixed, grid, wg, s, a)
                                : path_f64_gf<PDE, GEOM, STEP_ANY>(persistent, mixed, grid, wg, s, a);
}

template <int PDE>
static hipError_t path_f64(int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                          const PathKernelArgs &a) {
    if (a.pp.refl) {
        // reflecting sides: rectangles, one stream and k2 per launch, every step family
        if (mixed || geom != GeomRect) return hipErrorInvalidValue;
        if (!persistent)
            hipLaunchKernelGGL((plain_kernel<double, PDE, GeomRect, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
        else
            hipLaunchKernelGGL((path_kernel<double, PDE, false, GeomRect, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
        return hipGetLastError();
    }
    return geom == GeomDisc ? path_f64_g<PDE, GeomDisc>(persistent, mixed, grid, wg, s, a)
                            : path_f64_g<PDE, GeomRect>(persistent, mixed, grid, wg, s, a);
}

hipError_t launch_path_f64(int pde, int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                           const PathKernelArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return path_f64<PDE_MONEGROS>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_SINSIN: return path_f64<PDE_SINSIN>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF: return path_f64<PDE_VARCOEF>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_VARCOEF_LUT: return path_f64<PDE_VARCOEF_LUT>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_HELMHOLTZ: return path_f64<PDE_HELMHOLTZ>(geom, persistent, mixed, grid, wg, s, a);
        case PDE_TABLE: return path_f64<PDE_TABLE>(geom, persistent, mixed, grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

template <int GEOM> static int path_tiled_wg_g(int precision, int step) {
    if (step == STEP_CV)
        return precision == PDDS_FP64 ? tiled_wg<double, STEP_CV, GEOM>() : tiled_wg<float, STEP_CV, GEOM>();
    return precision == PDDS_FP64 ? tiled_wg<double, STEP_ANY, GEOM>() : tiled_wg<float, STEP_ANY, GEOM>();
}

int path_tiled_wg(int precision, int step, int geom) {
    return geom == GeomDisc ? path_tiled_wg_g<GeomDisc>(precision, step) : path_tiled_wg_g<GeomRect>(precision, step);
}

template <bool MIXED, int GEOM, int FAM>
static hipError_t path_tiled_f64_f(int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)path_kernel_tiled<double, MIXED, GEOM, FAM>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((path_kernel_tiled<double, MIXED, GEOM, FAM>), dim3(grid), dim3(tiled_wg<double, FAM, GEOM>()), lds, s, a);
    return hipGetLastError();
}

template <bool MIXED, int GEOM>
static hipError_t path_tiled_f64(int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    return a.pp.step == STEP_CV ? path_tiled_f64_f<MIXED, GEOM, STEP_CV>(grid, lds, s, a)
                                : path_tiled_f64_f<MIXED, GEOM, STEP_ANY>(grid, lds, s, a);
}

hipError_t launch_path_tiled_f64(int geom, bool mixed, int grid, size_t lds, hipStream_t s, const PathKernelArgs &a) {
    if (geom == GeomDisc)
        return mixed ? path_tiled_f64<true, GeomDisc>(grid, lds, s, a) : path_tiled_f64<false, GeomDisc>(grid, lds, s, a);
    return mixed ? path_tiled_f64<true, GeomRect>(grid, lds, s, a) : path_tiled_f64<false, GeomRect>(grid, lds, s, a);
}

template <int PDE>
static hipError_t paths_f64(int grid, int wg, hipStream_t s, const PathsArgs &a) {
    if (a.pp.refl)
        hipLaunchKernelGGL((paths_kernel<double, PDE, STEP_ANY, true>), dim3(grid), dim3(wg), 0, s, a);
    else if (a.pp.step == STEP_CV)
        hipLaunchKernelGGL((paths_kernel<double, PDE, STEP_CV>), dim3(grid), dim3(wg), 0, s, a);
    else
        hipLaunchKernelGGL((paths_kernel<double, PDE, STEP_ANY>), dim3(grid), dim3(wg), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_paths_f64(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a) {
    switch (pde) {
        case PDE_MONEGROS: return paths_f64<PDE_MONEGROS>(grid, wg, s, a);
        case PDE_SINSIN: return paths_f64<PDE_SINSIN>(grid, wg, s, a);
        case PDE_VARCOEF: return paths_f64<PDE_VARCOEF>(grid, wg, s, a);
        case PDE_VARCOEF_LUT: return paths_f64<PDE_VARCOEF_LUT>(grid, wg, s, a);
        case PDE_HELMHOLTZ: return paths_f64<PDE_HELMHOLTZ>(grid, wg, s, a);
        case PDE_TABLE: return paths_f64<PDE_TABLE>(grid, wg, s, a);
        default: return hipErrorInvalidValue;
    }
}

// One work-group per work item: per-side RBF moments per sub-batch, G_beta = -Psi^-T Phi per
// CSR column, sums in fixed order (sub-batch -> item), B sums per sub-batch.  A group of kNBG
// sub-batches is staged in LDS with, per (sub-batch, side), the list of its paths that exit
// through that side in path order (built with wave ballots, each exit's side read from its
// record's Y bits or found from its exit point, and counted per side).  A thread then owns one stencil
// entry e and, for each sub-batch b, adds Y phi_e(X) over the list of e's side in path order
// (the order of the per-exit loop it replaces), four exits per iteration so that four exp chains
// are in flight; Phi is kept [entry][sub-batch] so that the Psi^-T pass reads a row of kNBG
// values with two 16-byte LDS reads per pair of sub-batches.
// EQ: the launch carries reference values (ReduceArgs::wref) and also sums the knot equations
// SIDE (ReduceSide): how the exit side of a record is found
template <bool EQ, int SIDE>
__global__ void __launch_bounds__(kReduceWG) __attribute__((amdgpu_waves_per_eu(4))) reduce_kernel(ReduceArgs a) {
    extern __shared__ __align__(16) double smem[];
    __shared__ double s_bsb[kNBG], s_bbsb[kNBG];
    __shared__ double s_wsum[(kReduceWG / 64) * kNBG];
    __shared__ double s_exp2[64];
    __shared__ int s_lcnt[kNBG * 4];
    __shared__ unsigned int s_cnt[8];
    const int tid = threadIdx.x;
    const int WG = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = WG >> 6;
    const int64_t item = a.item_base + (int64_t)blockIdx.x;
    const ItemDesc it = a.items[item];
    const KnotDesc kd = a.knots[it.knot];
    const TableDesc tb = a.tables[kd.table];
    const StencilDesc &sd = tb.sten[kd.stencil];
    const int64_t g0 = kd.path_off + it.p_begin;
    const ExitRec *rr = a.rec.r + g0;
    const int S = a.S;
    const int GS = kNBG * S;
    double *Phi = smem;                                     // [ne_cap][kNBG]
    double *accS = smem + kNBG * a.ne_cap;
    double *accQ = accS + a.nnz_cap;
    double *gx = accQ + a.nnz_cap;
    double *gy = gx + GS;
    double *gY = gy + GS;
    double *gs = gY + GS;
    uint16_t *lst = (uint16_t *)(gs + GS);
    if (tid == 0)
        for (int q = 0; q < 8; ++q) s_cnt[q] = 0u;
    if (tid < 64) s_exp2[tid] = kExp2TabDev[tid];
    __syncthreads();
    // exit side of a record (stencil side 0..3, 4 global boundary, -1 none) and its Y
    auto side_of = [&](double x, double y, double Yr, double &Y) -> int {
        if constexpr (SIDE == SIDE_BITS) {
            const uint64_t yb = __builtin_bit_cast(uint64_t, Yr);
            Y = __builtin_bit_cast(double, yb & ~kSideBitsMask);
            return (int)(yb & kSideBitsMask) - 1;
        }
        Y = Yr;
        if constexpr (SIDE == SIDE_DISC) return classify_sector(sd.geo, a.gc[0], x, y);
        else return classify_coord_c(sd.rect, a.gc, x, y);
    };
    const int ne = sd.msz[0] + sd.msz[1] + sd.msz[2] + sd.msz[3];
    const int nnz = sd.csr_nnz;
    const int nsb = (it.n + S - 1) / S;
    // per-side sizes, first entries and Psi^-1 offsets as scalars picked by selects: a private
    // array indexed by a per-lane side would live in scratch memory
    const int ms0 = sd.msz[0], ms1 = sd.msz[1], ms2 = sd.msz[2], ms3 = sd.msz[3];
    const int64_t io0 = sd.ipsi_off[0], io1 = sd.ipsi_off[1], io2 = sd.ipsi_off[2], io3 = sd.ipsi_off[3];
    auto pick = [](int s, auto v0, auto v1, auto v2, auto v3) { return s == 0 ? v0 : s == 1 ? v1 : s == 2 ? v2 : v3; };
    auto msz_of = [&](int s) { return pick(s, ms0, ms1, ms2, ms3); };
    auto eoff_of = [&](int s) { return pick(s, 0, ms0, ms0 + ms1, ms0 + ms1 + ms2); };
    auto ipsi_of = [&](int s) { return pick(s, io0, io1, io2, io3); };
    for (int c = tid; c < nnz; c += WG) { accS[c] = 0.0; accQ[c] = 0.0; }
    double Bs = 0.0, BBs = 0.0, Rs = 0.0, RQs = 0.0;
    // (holding the next group's records in registers while a group is reduced, to overlap their
    // loads with the work, takes 8 more registers than the four-wave budget leaves: with the
    // exit side found below it spills and runs the reduction 35 % slower)
    for (int g = 0; g < nsb; g += kNBG) {
        const int nb = (nsb - g) < kNBG ? (nsb - g) : kNBG;
        const int q0 = g * S;                                   // first path of the group
        const int gn = (it.n - q0) < nb * S ? (it.n - q0) : nb * S;
        for (int p = tid; p < gn; p += WG) {
            const ExitRec rc = rr[q0 + p];
            gx[p] = rc.x;
            gy[p] = rc.y;
            gY[p] = rc.Y;
            gs[p] = rc.score;
        }
        __syncthreads();
        // per (sub-batch, side) lists in path order: a wave per sub-batch, 64 paths per ballot
        for (int b = wave; b < nb; b += nwaves) {
            const int p0 = b * S, p1 = (p0 + S) < gn ? (p0 + S) : gn;
            int cnt0 = 0, cnt1 = 0, cnt2 = 0, cnt3 = 0, cnt4 = 0;
            const unsigned long long below = (1ULL << lane) - 1ULL;
            for (int c0 = p0; c0 < p1; c0 += 64) {
                const int p = c0 + lane;
                // the exit side (FP32 records' Y loses its side bits here)
                int sdp = -2;
                if (p < p1) {
                    double Y;
                    sdp = side_of(gx[p], gy[p], gY[p], Y);
                    if constexpr (SIDE == SIDE_BITS) gY[p] = Y;
                }
                uint16_t *Lb = lst + (b * 4) * S;
                const unsigned long long m0 = __ballot(sdp == 0), m1 = __ballot(sdp == 1);
                const unsigned long long m2 = __ballot(sdp == 2), m3 = __ballot(sdp == 3);
                if (sdp == 0) Lb[0 * S + cnt0 + __popcll(m0 & below)] = (uint16_t)(p - p0);
                if (sdp == 1) Lb[1 * S + cnt1 + __popcll(m1 & below)] = (uint16_t)(p - p0);
                if (sdp == 2) Lb[2 * S + cnt2 + __popcll(m2 & below)] = (uint16_t)(p - p0);
                if (sdp == 3) Lb[3 * S + cnt3 + __popcll(m3 & below)] = (uint16_t)(p - p0);
                cnt0 += __popcll(m0); cnt1 += __popcll(m1); cnt2 += __popcll(m2); cnt3 += __popcll(m3);
                cnt4 += __popcll(__ballot(sdp == 4));
            }
            if (lane == 0) {
                s_lcnt[b * 4 + 0] = cnt0; s_lcnt[b * 4 + 1] = cnt1;
                s_lcnt[b * 4 + 2] = cnt2; s_lcnt[b * 4 + 3] = cnt3;
                // the item's exits per side
                atomicAdd(&s_cnt[0], (unsigned)cnt0); atomicAdd(&s_cnt[1], (unsigned)cnt1);
                atomicAdd(&s_cnt[2], (unsigned)cnt2); atomicAdd(&s_cnt[3], (unsigned)cnt3);
                atomicAdd(&s_cnt[4], (unsigned)cnt4);
                atomicAdd(&s_cnt[5], (unsigned)((p1 - p0) - (cnt0 + cnt1 + cnt2 + cnt3 + cnt4)));
            } else if (lane == 1) {
                // B sums of the sub-batch in path order
                double bs = 0.0, bb = 0.0;
                for (int p = p0; p < p1; ++p) { const double v = gs[p]; bs = bs + v; bb = __builtin_fma(v, v, bb); }
                s_bsb[b] = bs;
                s_bbsb[b] = bb;
            }
        }
        __syncthreads();
        // moments Phi[e][b] = sum over the exits of e's side in sub-batch b of Y phi_e(X)
        if (a.side_rbf) {
            // rectangle sides: a thread per (sub-batch, side, chunk of kRbfChunk entries) walks the
            // side's exits of the sub-batch in path order, forms the chunk's RBF values of each
            // (rbf_side_chunk) and adds Y phi into one register sum per entry -- per entry the
            // order of the direct form
            const int cb1 = (ms0 + kRbfChunk - 1) / kRbfChunk;
            const int cb2 = cb1 + (ms1 + kRbfChunk - 1) / kRbfChunk;
            const int cb3 = cb2 + (ms2 + kRbfChunk - 1) / kRbfChunk;
            const int cb4 = cb3 + (ms3 + kRbfChunk - 1) / kRbfChunk;
            for (int t = tid; t < nb * cb4; t += WG) {
                const int b = t / cb4, r = t - b * cb4;
                const int s = r < cb1 ? 0 : r < cb2 ? 1 : r < cb3 ? 2 : 3;
                const int m = msz_of(s);
                const int j0 = (r - pick(s, 0, cb1, cb2, cb3)) * kRbfChunk;
                const int len = (m - j0) < kRbfChunk ? (m - j0) : kRbfChunk;
                const int e0 = sd.ent_off + eoff_of(s);
                const double *px = tb.entx + e0, *py = tb.enty + e0;
                const double ic2 = tb.entic2[e0];
                double sx = 0.0, sy = 0.0, ss = 1.0;
                if (m > 1) {
                    sx = px[1] - px[0];
                    sy = py[1] - py[0];
                    ss = sx * sx + sy * sy;
                }
                const double q = dexp_tab_nonpos(-(2.0 * ss) * ic2, s_exp2);
                const uint16_t *L = lst + (b * 4 + s) * S;
                const int nl = s_lcnt[b * 4 + s];
                const int pb = b * S;
                double acc[kRbfChunk];
#pragma unroll
                for (int k = 0; k < kRbfChunk; ++k) acc[k] = 0.0;
                for (int l = 0; l < nl; ++l) {
                    const int p = pb + L[l];
                    rbf_side_chunk(px, py, j0, sx, sy, ss, q, gx[p], gy[p], ic2, s_exp2, gY[p], acc);
                }
#pragma unroll
                for (int k = 0; k < kRbfChunk; ++k)
                    if (k < len) Phi[(eoff_of(s) + j0 + k) * kNBG + b] = acc[k];
            }
        } else
        for (int e = tid; e < ne; e += WG) {
            const int se = tb.ents[sd.ent_off + e];
            const double pxe = tb.entx[sd.ent_off + e], pye = tb.enty[sd.ent_off + e];
            const double ic2 = tb.entic2[sd.ent_off + e];
            for (int b = 0; b < nb; ++b) {
                const uint16_t *L = lst + (b * 4 + se) * S;
                const int nl = s_lcnt[b * 4 + se];
                const int pb = b * S;
                double acc = 0.0;
                int l = 0;
                for (; l + 3 < nl; l += 4) {
                    const int pa = pb + L[l], pc = pb + L[l + 1], pd = pb + L[l + 2], pe = pb + L[l + 3];
                    const double fa = rbf_tab(pxe, pye, gx[pa], gy[pa], ic2, s_exp2);
                    const double fc = rbf_tab(pxe, pye, gx[pc], gy[pc], ic2, s_exp2);
                    const double fd = rbf_tab(pxe, pye, gx[pd], gy[pd], ic2, s_exp2);
                    const double fe = rbf_tab(pxe, pye, gx[pe], gy[pe], ic2, s_exp2);
                    acc = __builtin_fma(gY[pa], fa, acc);
                    acc = __builtin_fma(gY[pc], fc, acc);
                    acc = __builtin_fma(gY[pd], fd, acc);
                    acc = __builtin_fma(gY[pe], fe, acc);
                }
                for (; l < nl; ++l) {
                    const int pa = pb + L[l];
                    acc = __builtin_fma(gY[pa], rbf_tab(pxe, pye, gx[pa], gy[pa], ic2, s_exp2), acc);
                }
                Phi[e * kNBG + b] = acc;
            }
        }
        __syncthreads();
        double tp[kNBG];   // this thread's columns' sum of G_b,c u_ref (equation residual)
#pragma unroll
        for (int b = 0; b < kNBG; ++b) tp[b] = 0.0;
        for (int c = tid; c < nnz; c += WG) {
            double G[kNBG];
#pragma unroll
            for (int b = 0; b < kNBG; ++b) G[b] = 0.0;
            const int ec[2] = {tb.e1[sd.csr_off + c], tb.e2[sd.csr_off + c]};
            for (int q = 0; q < 2; ++q) {
                const int e = ec[q];
                if (e < 0) continue;
                const int se = tb.ents[sd.ent_off + e];
                const int m = msz_of(se);
                const int es = eoff_of(se);
                const int el = e - es;
                const double *A = tb.ipsi + ipsi_of(se) + el;
                const double2 *ph = (const double2 *)(Phi + es * kNBG);
                double H[kNBG];
#pragma unroll
                for (int b = 0; b < kNBG; ++b) H[b] = 0.0;
                // Psi^-1 column el, eight rows per batch of loads so that their latency overlaps
                int j = 0;
                for (; j + 8 <= m; j += 8) {
                    double av[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) av[u] = A[(int64_t)(j + u) * m];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
#pragma unroll
                        for (int b2 = 0; b2 < kNBG / 2; ++b2) {
                            const double2 v = ph[(j + u) * (kNBG / 2) + b2];
                            H[2 * b2] = __builtin_fma(av[u], v.x, H[2 * b2]);
                            H[2 * b2 + 1] = __builtin_fma(av[u], v.y, H[2 * b2 + 1]);
                        }
                    }
                }
                for (; j < m; ++j) {
                    const double aj = A[(int64_t)j * m];
#pragma unroll
                    for (int b2 = 0; b2 < kNBG / 2; ++b2) {
                        const double2 v = ph[j * (kNBG / 2) + b2];
                        H[2 * b2] = __builtin_fma(aj, v.x, H[2 * b2]);
                        H[2 * b2 + 1] = __builtin_fma(aj, v.y, H[2 * b2 + 1]);
                    }
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
            if constexpr (EQ) {
                const double wc = a.wref[kd.out_off + c];
#pragma unroll
                for (int b = 0; b < kNBG; ++b) tp[b] = __builtin_fma(G[b], wc, tp[b]);
            }
        }
        if constexpr (EQ) {
            // sum over the columns: a butterfly over each wave, then the waves in order
#pragma unroll
            for (int b = 0; b < kNBG; ++b) {
                double v = tp[b];
                for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
                if (lane == 0) s_wsum[wave * kNBG + b] = v;
            }
            __syncthreads();
        }
        if (tid == 0) {
            for (int b = 0; b < nb; ++b) { Bs = Bs + s_bsb[b]; BBs = BBs + s_bbsb[b]; }
            if constexpr (EQ)
                for (int b = 0; b < nb; ++b) {
                    double T = 0.0;
                    for (int w = 0; w < nwaves; ++w) T = T + s_wsum[w * kNBG + b];
                    const int p0 = (g + b) * S;
                    const int p1 = (p0 + S) < it.n ? (p0 + S) : it.n;
                    const double R = s_bsb[b] - T;
                    Rs = Rs + R;
                    RQs = RQs + (R * R) / (double)(p1 - p0);
                }
        }
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
        st.Rs = Rs;
        st.RQs = RQs;
        for (int q = 0; q < 6; ++q) st.cnt[q] = s_cnt[q];
        a.stats[item] = st;
    }
}

template <bool EQ, int SIDE> static hipError_t launch_reduce_t(int grid, int wg, size_t shmem, hipStream_t s,
                                                            const ReduceArgs &a) {
    if (shmem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)reduce_kernel<EQ, SIDE>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((reduce_kernel<EQ, SIDE>), dim3(grid), dim3(wg), shmem, s, a);
    return hipGetLastError();
}

template <bool EQ> static hipError_t launch_reduce_eq(int grid, int wg, size_t shmem, hipStream_t s, const ReduceArgs &a) {
    switch (a.side_mode) {
        case SIDE_BITS: return launch_reduce_t<EQ, SIDE_BITS>(grid, wg, shmem, s, a);
        case SIDE_RECT: return launch_reduce_t<EQ, SIDE_RECT>(grid, wg, shmem, s, a);
        case SIDE_DISC: return launch_reduce_t<EQ, SIDE_DISC>(grid, wg, shmem, s, a);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce(int grid, int wg, size_t shmem, hipStream_t s, const ReduceArgs &a) {
    return a.wref ? launch_reduce_eq<true>(grid, wg, shmem, s, a) : launch_reduce_eq<false>(grid, wg, shmem, s, a);
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
        double Bs = 0.0, BBs = 0.0, Rs = 0.0, RQs = 0.0;
        KnotStat ks;
        ks.steps = f.kacc[2 * (int64_t)k];
        ks.capped = f.kacc[2 * (int64_t)k + 1];
        for (int q = 0; q < 6; ++q) ks.cnt[q] = 0;
        for (int64_t w = w0; w < w1; ++w) {
            const ItemStat st = f.stats[w];
            Bs = Bs + st.Bs;
            BBs = BBs + st.BBs;
            Rs = Rs + st.Rs;
            RQs = RQs + st.RQs;
            for (int q = 0; q < 6; ++q) ks.cnt[q] += st.cnt[q];
        }
        ks.B = Bs / N;
        ks.BB = BBs / N;
        ks.Rm = Rs / N;
        ks.Rv = nb > 1 ? (RQs - N * ks.Rm * ks.Rm) / (double)(nb - 1) : 0.0;
        f.kstats[k] = ks;
    }
}

hipError_t launch_finalize(int grid, int wg, hipStream_t s, const FinalizeArgs &f) {
    hipLaunchKernelGGL(finalize_kernel, dim3(grid), dim3(wg), 0, s, f);
    return hipGetLastError();
}

// Sets n 64-bit words to v: the path queues' counters are cleared by a kernel rather than a
// memset node, so the sequence stays capturable in a HIP graph.
__global__ void fill_u64_kernel(unsigned long long *p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int wg = 256;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + wg - 1) / wg)), dim3(wg), 0, s, p, n, v);
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
