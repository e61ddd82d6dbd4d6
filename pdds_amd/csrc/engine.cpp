// engine.cpp -- C ABI of the path engine (include/pdds_engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdds_engine.h"
#include "engine_types.h"
#include "mesh.h"
#include "solver.h"
#include "engine_internal.h"

using namespace pdds;

uint64_t next_mesh_id() {
    static std::mutex mu;
    static uint64_t id = 0;
    std::lock_guard<std::mutex> lk(mu);
    return ++id;
}

pdds_status fail(pdds_engine *e, pdds_status s, const std::string &msg) {
    if (e) e->err = msg;
    return s;
}

// ---------------------------------------------------------------------------
// No-allocation scratch: one fixed device array per device, held from the library's load until
// the process ends.  In host code `scratch` names a host-side stand-in, not device memory; its
// device address on each device comes from hipGetSymbolAddress (noalloc_scratch_of), and only
// that address is ever handed to kernels or copies.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PDDS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_call_mu[kMaxDevices];   // serialises the calls that use a device's array
}  // namespace

// Device address of `device`'s scratch array, looked up once per device with that device
// current (the caller has made it current).
static hipError_t noalloc_scratch_of(int device, char **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess) return r;
        g_scratch_ptr[device] = (char *)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// device buffer for slot i: grow-only and kept by the engine, or, while the engine has an arena
// set, the next piece of the arena
pdds_status dbuf(pdds_engine *e, int slot, size_t bytes, void **out) {
    if (bytes == 0) bytes = 8;
    if (e->arena) {
        Arena &a = *e->arena;
        const size_t off = (a.used + 255) & ~(size_t)255;
        if (off > a.cap || bytes > a.cap - off)
            return fail(e, PDDS_ERR_FULL, "the call needs more device memory than the " +
                                              std::to_string(a.cap >> 20) + " MiB no-allocation scratch");
        *out = a.base + off;
        a.used = off + bytes;
        return PDDS_OK;
    }
    DevBuf &b = e->bufs[slot];
    if (b.n < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        if (hipMalloc(&b.p, bytes) != hipSuccess) return fail(e, PDDS_ERR_NOMEM, "device allocation failed");
        b.n = bytes;
    }
    *out = b.p;
    return PDDS_OK;
}

bool valid_pde(int pde) { return pde >= PDDS_PDE_MONEGROS && pde <= PDDS_PDE_TABLE; }

pdds_status make_params(pdds_engine *e, const pdds_problem *p, const pdds_run *r, PdeParams &pp) {
    if (!p || !valid_pde(p->pde)) return fail(e, PDDS_ERR_INVALID, "invalid problem descriptor");
    std::memset(&pp, 0, sizeof(pp));
    pp.k2 = p->k2;
    pp.step = r ? r->step : STEP_CV;
    // at least one side absorbs, so that every stopping domain (a stencil rectangle, or the whole
    // rectangle of a point evaluation) has a side where paths stop
    if (p->refl < 0 || p->refl > 14 || !(p->robin <= 0.0) || !std::isfinite(p->robin) || !(p->rho >= 0.0))
        return fail(e, PDDS_ERR_INVALID, "reflecting sides: refl in 0..14 (one side at least absorbs), robin <= 0, rho >= 0");
    pp.refl = p->refl;
    pp.robin = p->robin;
    pp.rho = p->rho;
    if (p->pde == PDDS_PDE_VARCOEF_LUT) {
        if (!p->lut || p->lut->nf != 5) return fail(e, PDDS_ERR_INVALID, "PDDS_PDE_VARCOEF_LUT needs a LUT with 5 fields");
        pp.lut = p->lut->view;
    }
    if (p->pde == PDDS_PDE_TABLE) {
        if (!p->lut || p->lut->nf != 8 || !p->ulut || p->ulut->nf != 1)
            return fail(e, PDDS_ERR_INVALID, "PDDS_PDE_TABLE needs an 8-field lut and a 1-field ulut");
        pp.lut = p->lut->view;
        pp.ulut = p->ulut->view;
    }
    if (pp.step == STEP_VRCV) {
        if (!p->vlut || p->vlut->nf != 4) return fail(e, PDDS_ERR_INVALID, "PDDS_STEP_VRCV needs a 4-field vlut");
        pp.vlut = p->vlut->view;
    }
    return PDDS_OK;
}

// host evaluation of the Dirichlet data (boundary knots: B = g(X0), Solvers/EMFKACSolver.hpp:720-727);
// a tabulated u is read from the host copy of its table
double host_g(const pdds_problem *p, double x, double y, const PdeParams &pp) {
    switch (p->pde) {
        case PDE_MONEGROS: return pde_g<PDE_MONEGROS, double>(x, y, pp);
        case PDE_SINSIN: return pde_g<PDE_SINSIN, double>(x, y, pp);
        case PDE_HELMHOLTZ: return pde_g<PDE_HELMHOLTZ, double>(x, y, pp);
        case PDE_TABLE: {
            PdeParams hp = pp;
            hp.ulut = p->ulut->hview;
            return pde_g<PDE_TABLE, double>(x, y, hp);
        }
        default: return pde_g<PDE_VARCOEF, double>(x, y, pp);
    }
}

// sigma at (x, y) on the host (tabulated coefficients from the host copies of their tables)
static void host_sigma(const pdds_problem *p, const PdeParams &pp, double x, double y, double s[4]) {
    PdeParams hp = pp;
    if (p->lut) hp.lut = p->lut->hview;
    if (p->ulut) hp.ulut = p->ulut->hview;
    Coef<double> co;
    switch (p->pde) {
        case PDE_MONEGROS: pde_coef<PDE_MONEGROS, double>(x, y, hp, co); break;
        case PDE_SINSIN: pde_coef<PDE_SINSIN, double>(x, y, hp, co); break;
        case PDE_VARCOEF: pde_coef<PDE_VARCOEF, double>(x, y, hp, co); break;
        case PDE_VARCOEF_LUT: pde_coef<PDE_VARCOEF_LUT, double>(x, y, hp, co); break;
        case PDE_HELMHOLTZ: pde_coef<PDE_HELMHOLTZ, double>(x, y, hp, co); break;
        default: pde_coef<PDE_TABLE, double>(x, y, hp, co); break;
    }
    s[0] = co.s00; s[1] = co.s01; s[2] = co.s10; s[3] = co.s11;
}

// Step of a knot's paths: the run's h, or with adaptive_h the distance rule of
// PDDAlgorithms/PDDSparseGM.cpp:4025-4029 -- h_k = min(h, (d / (2 |beta| |N^T sigma(X0)|))^2) for
// a start point at signed distance d <= -1e-4 from its stopping domain (N the nearest side's
// outward normal), so that the Gobet shift |beta| |N^T sigma| sqrt(h_k) is at most d / 2
double knot_step(const pdds_problem *p, const PdeParams &pp, const pdds_run *r, int geom, const double rect[4],
                 const double geo[6], double x0, double y0) {
    if (!r->adaptive_h) return r->h;
    double ex, ey, nx, ny;
    int kind;
    const double d = geom == GeomDisc ? sector_dist<double>(geo, x0, y0, ex, ey, nx, ny, kind)
                                      : rect_dist<double>(rect, x0, y0, ex, ey, nx, ny, kind);
    if (!(d <= -1e-4)) return r->h;
    double s[4];
    host_sigma(p, pp, x0, y0, s);
    const double ga = nx * s[0] + ny * s[2], gb = nx * s[1] + ny * s[3];
    const double q = d / (std::sqrt(ga * ga + gb * gb) * 2.0 * std::fabs(r->beta));
    return std::min(r->h, q * q);
}

pdds_status check_run(pdds_engine *e, const pdds_run *r) {
    if (!r) return fail(e, PDDS_ERR_INVALID, "null run parameters");
    if (!(r->h > 0.0) || !std::isfinite(r->h)) return fail(e, PDDS_ERR_INVALID, "h must be positive");
    if (r->precision != PDDS_FP64 && r->precision != PDDS_FP32) return fail(e, PDDS_ERR_INVALID, "bad precision");
    if (r->wg_size != 0 && r->wg_size != 64 && r->wg_size != 128 && r->wg_size != 256)
        return fail(e, PDDS_ERR_INVALID, "wg_size must be 64, 128 or 256");
    if (r->sub_batch < 0 || r->item_paths < 0 || r->grid_wgs < 0 || r->max_steps < 0)
        return fail(e, PDDS_ERR_INVALID, "negative launch parameter");
    if (r->paths_per_lane < 0 || r->paths_per_lane > 64)
        return fail(e, PDDS_ERR_INVALID, "paths_per_lane must be in [0, 64]");
    static_assert(PDDS_STEP_CV == STEP_CV && PDDS_STEP_PLAIN == STEP_PLAIN && PDDS_STEP_VR == STEP_VR &&
                      PDDS_STEP_VRCV == STEP_VRCV && PDDS_PDE_TABLE == PDE_TABLE,
                  "step family and problem values");
    if (r->step < PDDS_STEP_CV || r->step > PDDS_STEP_VRCV) return fail(e, PDDS_ERR_INVALID, "bad step family");
    if (r->adaptive_h != 0 && r->adaptive_h != 1) return fail(e, PDDS_ERR_INVALID, "adaptive_h must be 0 or 1");
    if (r->lut_pad < 0 || r->lut_pad > 15) return fail(e, PDDS_ERR_INVALID, "lut_pad must be in [0, 15]");
    return PDDS_OK;
}

// Paths per work item of the reduction: a multiple of S, about N/16, at most 8192.
int choose_P(int64_t n, int S) {
    int64_t q = (n + 16 * (int64_t)S - 1) / (16 * (int64_t)S);
    if (q < 1) q = 1;
    if (q * S > 8192) q = 8192 / S;
    if (q < 1) q = 1;
    return (int)(q * S);
}

template <class T>
pdds_status dalloc_copy(pdds_engine *e, const std::vector<T> &v, T **out, size_t &bytes) {
    size_t nb = std::max<size_t>(v.size() * sizeof(T), 8);
    if (hipMalloc((void **)out, nb) != hipSuccess) return fail(e, PDDS_ERR_NOMEM, "device allocation failed");
    bytes += nb;
    if (!v.empty()) HIPCHK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return PDDS_OK;
}

// Upload host stencil tables into T (device memory owned by T).
pdds_status build_tables(pdds_engine *e, const std::vector<StencilDesc> &sten, const std::vector<double> &entx,
                         const std::vector<double> &enty, const std::vector<int8_t> &ents,
                         const std::vector<double> &entic2,
                         const std::vector<int32_t> &e1, const std::vector<int32_t> &e2,
                         const std::vector<double> &ipsi, DevTables &T) {
    T.release();
    pdds_status st;
    if ((st = dalloc_copy(e, sten, &T.sten, T.bytes)) || (st = dalloc_copy(e, entx, &T.entx, T.bytes)) ||
        (st = dalloc_copy(e, enty, &T.enty, T.bytes)) || (st = dalloc_copy(e, ents, &T.ents, T.bytes)) ||
        (st = dalloc_copy(e, entic2, &T.entic2, T.bytes)) ||
        (st = dalloc_copy(e, e1, &T.e1, T.bytes)) || (st = dalloc_copy(e, e2, &T.e2, T.bytes)) ||
        (st = dalloc_copy(e, ipsi, &T.ipsi, T.bytes))) {
        T.release();
        return st;
    }
    T.host_sten = sten;
    T.n_sten = (int)sten.size();
    return PDDS_OK;
}

// Plane-a texels a path inside rectangle rect can touch, one texel of slack on each side for
// the rounding of the coordinate map: padded columns [c0, c0 + tw), rows [r0, r0 + th).
void lut_tile_of(const LutView &L, const double rect[4], int &c0, int &r0, int &tw, int &th) {
    auto span = [](double a, double b, double o, double sc, int n, int np, int &lo, int &len) {
        auto idx = [&](double x) {
            double t = (x - o) * sc;
            t = std::min(std::max(t, -2.0), (double)n + 1.0);
            return (int)std::floor(t);
        };
        const int i0 = std::max(idx(a) - 1, -2), i1 = std::min(idx(b) + 1, n + 1);
        lo = std::max(i0 + kLutPad - 1, 0);
        const int hi = std::min(i1 + kLutPad + 2, np - 1);
        len = hi - lo + 1;
    };
    const int py = L.ny + 2 * kLutPad + 1;
    span(rect[0], rect[2], L.ox, L.sx, L.nx, L.px, c0, tw);
    span(rect[1], rect[3], L.oy, L.sy, L.ny, py, r0, th);
}

pdds_status run_assembly(pdds_engine *e, const pdds_problem *prob, const pdds_run *r, const double glob[4],
                         const std::vector<const DevTables *> &tabs, const std::vector<AsmKnot> &knots,
                         int64_t n_vals, double *G_host, double *Gvar_host, AsmOut &out, const double *W_host) {
    auto host_sd = [&](const AsmKnot &k) -> const StencilDesc & { return tabs[k.table]->host_sten[k.stencil]; };
    PdeParams pp;
    pdds_status st = make_params(e, prob, r, pp);
    if (st) return st;
    const int S = r->sub_batch > 0 ? r->sub_batch : 32;
    const int wg = r->wg_size > 0 ? r->wg_size : 256;
    const bool persistent = r->plain == 0;
    const int64_t max_steps = r->max_steps > 0 ? r->max_steps : (int64_t)1 << 26;
    if (r->item_paths > 0 && r->item_paths % S != 0)
        return fail(e, PDDS_ERR_INVALID, "item_paths must be a multiple of sub_batch");
    // a path's step count is a 32-bit register (the two-phase form adds a second cap of steps)
    if (max_steps >= ((int64_t)1 << 27)) return fail(e, PDDS_ERR_INVALID, "max_steps must be below 2^27");
    out.S = S;
    out.wg = wg;
    out.kstats.assign(knots.size(), KnotStat{});
    if (knots.empty()) return PDDS_OK;
    const int geom = tabs[0]->geom;
    for (const DevTables *t : tabs)
        if (t->geom != geom) return fail(e, PDDS_ERR_INVALID, "one assembly mixes stencil geometries");
    if (pp.refl && geom != GeomRect) return fail(e, PDDS_ERR_UNSUPPORTED, "reflecting sides need rectangular stencils");
    // LDS-tiled form (coefficient LUT with a fifth-field quad plane, persistent launch): knots
    // are grouped into patches of equal tile, tiles that fit in LDS first, in first-seen order;
    // not with reflecting sides, whose steps run in the persistent and plain kernels
    const bool tiled =
        persistent && prob->pde == PDDS_PDE_VARCOEF_LUT && r->lut_lds != 0 && pp.lut.nf == 5 && pp.refl == 0;
    const size_t tile_cap = e->lds_max > 1024 ? e->lds_max - 1024 : 0;
    std::vector<size_t> ord(knots.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::vector<std::array<int, 5>> tkey(knots.size());   // c0, r0, tw, th, fits
    if (tiled) {
        std::map<std::array<int, 4>, int> first;
        std::vector<int> rank(knots.size());
        for (size_t i = 0; i < knots.size(); ++i) {
            int c0, r0, tw, th;
            lut_tile_of(pp.lut, host_sd(knots[i]).bbox, c0, r0, tw, th);
            const int fits = (size_t)(tw + r->lut_pad) * th * sizeof(float4) <= tile_cap ? 1 : 0;
            tkey[i] = {c0, r0, tw, th, fits};
            const std::array<int, 4> k4 = fits ? std::array<int, 4>{c0, r0, tw, th} : std::array<int, 4>{-1, -1, -1, -1};
            auto it = first.find(k4);
            if (it == first.end()) it = first.emplace(k4, fits ? (int)i : (int)knots.size() + (int)i).first;
            rank[i] = it->second;
        }
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return rank[a] < rank[b]; });
        out.wg = path_tiled_wg(r->precision, r->step, geom);
    }
    // knots are processed in batches whose exit records fit the record budget
    const int64_t rec_bytes = (int64_t)sizeof(ExitRec);
    const int64_t budget = e->record_budget / rec_bytes;
    void *p;
    double *dG, *dGv;
    if ((st = dbuf(e, B_G, (size_t)std::max<int64_t>(n_vals, 1) * 8, &p))) return st; dG = (double *)p;
    if ((st = dbuf(e, B_GVAR, (size_t)std::max<int64_t>(n_vals, 1) * 8, &p))) return st; dGv = (double *)p;
    double *dW = nullptr;
    if (W_host && n_vals > 0) {
        if ((st = dbuf(e, B_WREF, (size_t)n_vals * 8, &p))) return st;
        dW = (double *)p;
        HIPCHK(hipMemcpyAsync(dW, W_host, (size_t)n_vals * 8, hipMemcpyHostToDevice, e->stream));
    }
    std::vector<TableDesc> tdesc(tabs.size());
    for (size_t t = 0; t < tabs.size(); ++t)
        tdesc[t] = TableDesc{tabs[t]->sten, tabs[t]->entx, tabs[t]->enty, tabs[t]->ents, tabs[t]->entic2,
                             tabs[t]->e1, tabs[t]->e2, tabs[t]->ipsi};
    TableDesc *dtab;
    if ((st = upload(e, B_TABLES, tdesc, &dtab))) return st;
    double kernel_ms = 0.0, reduce_ms = 0.0;
    int grid_used = 0, Pmax_all = 0;
    int64_t items_all = 0;
    size_t k0 = 0;
    while (k0 < knots.size()) {
        // batch [k0, k1)
        size_t k1 = k0;
        int64_t tot = 0;
        while (k1 < knots.size() && (k1 == k0 || tot + knots[ord[k1]].n <= budget)) tot += knots[ord[k1++]].n;
        const int nk = (int)(k1 - k0);
        if (tot >= ((int64_t)1 << 32)) return fail(e, PDDS_ERR_UNSUPPORTED, "a knot with 2^32 or more paths");
        std::vector<KnotDesc> kd(nk);
        std::vector<ItemDesc> items;
        int64_t part = 0, off = 0;
        int ne_cap = 1, nnz_cap = 1;
        for (int q = 0; q < nk; ++q) {
            const AsmKnot &k = knots[ord[k0 + q]];
            const StencilDesc &sd = host_sd(k);
            const int P = r->item_paths > 0 ? r->item_paths : choose_P(k.n, S);
            Pmax_all = std::max(Pmax_all, P);
            ne_cap = std::max(ne_cap, sd.msz[0] + sd.msz[1] + sd.msz[2] + sd.msz[3]);
            nnz_cap = std::max(nnz_cap, sd.csr_nnz);
            kd[q].x0 = k.x0; kd[q].y0 = k.y0; kd[q].knot_id = k.id; kd[q].stencil = k.stencil;
            kd[q].table = k.table; kd[q].stream = k.has_stream ? k.stream : r->stream;
            kd[q].n_paths = k.n; kd[q].path_off = off; kd[q].out_off = k.out_off;
            kd[q].k2 = k.has_k2 ? k.k2 : pp.k2;
            kd[q].h = k.h > 0.0 ? k.h : knot_step(prob, pp, r, geom, sd.rect, sd.geo, k.x0, k.y0);
            kd[q].sqrth = std::sqrt(kd[q].h);
            kd[q].item_begin = (int64_t)items.size();
            for (int64_t p0 = 0; p0 < k.n; p0 += P) {
                ItemDesc it;
                it.knot = q;
                it.n = (int32_t)std::min<int64_t>(P, k.n - p0);
                it.p_begin = p0;
                it.part_off = part;
                part += 2 * (int64_t)sd.csr_nnz;
                items.push_back(it);
            }
            kd[q].item_count = (int64_t)items.size() - kd[q].item_begin;
            off += k.n;
        }
        items_all += (int64_t)items.size();
        // knots of other streams, k2 or h than the run's (a serving round, adaptive h): per-lane values
        bool mixed = false;
        for (int q = 0; q < nk; ++q)
            mixed = mixed || kd[q].stream != r->stream || kd[q].k2 != pp.k2 || kd[q].h != r->h;
        if (mixed && pp.refl)
            return fail(e, PDDS_ERR_UNSUPPORTED, "reflecting sides with per-knot streams, k2 or h in one launch");
        const size_t shmem = reduce_lds_bytes(ne_cap, nnz_cap, S);
        if (shmem > 160 * 1024 || S > 65535)
            return fail(e, PDDS_ERR_UNSUPPORTED, "stencil or sub-batch too large for the LDS reduction buffers");
        KnotDesc *dk; ItemDesc *dit;
        if ((st = upload(e, B_KNOTS, kd, &dk))) return st;
        if ((st = upload(e, B_ITEMS, items, &dit))) return st;
        Records rec;
        if ((st = dbuf(e, B_REC, (size_t)tot * sizeof(ExitRec), &p))) return st; rec.r = (ExitRec *)p;
        if ((st = dbuf(e, B_KACC, (size_t)nk * 2 * sizeof(unsigned long long), &p))) return st;
        rec.kacc = (unsigned long long *)p;
        HIPCHK(launch_fill_u64(rec.kacc, 2 * (int64_t)nk, 0ULL, e->stream));
        ItemStat *dstats; double *dpart; KnotStat *dks; unsigned long long *dctr;
        if ((st = dbuf(e, B_STATS, items.size() * sizeof(ItemStat), &p))) return st; dstats = (ItemStat *)p;
        if ((st = dbuf(e, B_PART, (size_t)std::max<int64_t>(part, 1) * sizeof(double), &p))) return st; dpart = (double *)p;
        if ((st = dbuf(e, B_KSTAT, (size_t)nk * sizeof(KnotStat), &p))) return st; dks = (KnotStat *)p;
        // patches of the batch: runs of equal tile among the ordered knots
        std::vector<PatchDesc> patches;
        size_t lds_bytes = 0;
        if (tiled) {
            int64_t poff = 0;
            for (int q = 0; q < nk; ++q) {
                const std::array<int, 5> &tk = tkey[ord[k0 + q]];
                const bool same = !patches.empty() &&
                                  (tk[4] ? (patches.back().tiled && patches.back().c0 == tk[0] && patches.back().r0 == tk[1] &&
                                            patches.back().tw == tk[2] && patches.back().th == tk[3])
                                         : !patches.back().tiled);
                if (!same) {
                    PatchDesc pd;
                    std::memset(&pd, 0, sizeof(pd));
                    pd.pbeg = poff;
                    pd.tiled = tk[4];
                    if (tk[4]) {
                        pd.c0 = tk[0]; pd.r0 = tk[1]; pd.tw = tk[2]; pd.th = tk[3];
                        pd.pitch = tk[2] + r->lut_pad;
                        lds_bytes = std::max(lds_bytes, (size_t)pd.pitch * tk[3] * sizeof(float4));
                    }
                    patches.push_back(pd);
                }
                poff += kd[q].n_paths;
                patches.back().pend = poff;
            }
        }
        const size_t nctr = std::max<size_t>(kMaxQueues, patches.size());
        PatchDesc *dpatch = nullptr;
        if (tiled && (st = upload(e, B_PATCH, patches, &dpatch))) return st;
        if ((st = dbuf(e, B_COUNTER, nctr * sizeof(unsigned long long), &p))) return st; dctr = (unsigned long long *)p;
        HIPCHK(launch_fill_u64(dctr, (int64_t)nctr, 0ULL, e->stream));
        // path kernel
        PathKernelArgs pa;
        std::memset(&pa, 0, sizeof(pa));
        pa.knots = dk; pa.n_knots = nk; pa.tables = dtab; pa.total = tot; pa.base = 0; pa.counter = dctr; pa.rec = rec;
        for (int i = 0; i < 4; ++i) pa.glob[i] = glob[i];
        pa.h = r->h; pa.sqrth = std::sqrt(r->h); pa.beta = r->beta;
        pa.seed = r->seed; pa.stream = r->stream; pa.max_steps = max_steps; pa.pp = pp;
        pa.pool_chunk = 64 * (int64_t)(r->paths_per_lane > 0 ? r->paths_per_lane : 4);
        int grid;
        if (persistent) {
            grid = r->grid_wgs > 0 ? r->grid_wgs : e->path_wgs_per_cu * e->n_cu;
            grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (tot + wg - 1) / wg));
        } else {
            grid = (int)std::min<int64_t>((tot + wg - 1) / wg, (int64_t)1 << 30);
        }
        pa.n_queues = persistent ? std::min(kMaxQueues, grid) : 1;
        pa.patches = dpatch;
        pa.n_patches = (int32_t)patches.size();
        const int twg = path_tiled_wg(r->precision, r->step, geom);
        if (tiled) {
            grid = r->grid_wgs > 0 ? r->grid_wgs
                                   : (int)std::max<int64_t>(1, std::min<int64_t>(e->n_cu, (tot + 8 * twg - 1) / (8 * twg)));
        }
        grid_used = std::max(grid_used, grid);
        HIPCHK(hipEventRecord(e->ev[0], e->stream));
        const int64_t per_launch = persistent ? tot : (int64_t)grid * wg;
        for (int64_t base = 0; base < tot; base += per_launch) {
            pa.base = base;
            const int g = persistent ? grid : (int)std::min<int64_t>(grid, (tot - base + wg - 1) / wg);
            hipError_t lr;
            if (tiled)
                lr = r->precision == PDDS_FP64 ? launch_path_tiled_f64(geom, mixed, g, lds_bytes, e->stream, pa)
                                               : launch_path_tiled_f32(geom, mixed, g, lds_bytes, e->stream, pa);
            else
                lr = r->precision == PDDS_FP64 ? launch_path_f64(prob->pde, geom, persistent, mixed, g, wg, e->stream, pa)
                                               : launch_path_f32(prob->pde, geom, persistent, mixed, g, wg, e->stream, pa);
            if (lr != hipSuccess) return fail(e, PDDS_ERR_DEVICE, std::string("path launch: ") + hipGetErrorString(lr));
        }
        HIPCHK(hipEventRecord(e->ev[1], e->stream));
        // reduction + finalize
        ReduceArgs ra;
        std::memset(&ra, 0, sizeof(ra));
        ra.knots = dk; ra.tables = dtab; ra.items = dit; ra.rec = rec;
        ra.stats = dstats; ra.partials = dpart; ra.ne_cap = ne_cap; ra.nnz_cap = nnz_cap; ra.S = S;
        ra.wref = dW;
        ra.side_rbf = geom == GeomRect ? 1 : 0;   // rectangle sides: the chunked recurrence
        // exit sides: FP32 records carry them; FP64 exits are classified as write_exit does, on
        // the global rectangle's centre and half-sides (rect_dist's) or the disc's radius
        ra.side_mode = r->precision != PDDS_FP64 ? SIDE_BITS : geom == GeomDisc ? SIDE_DISC : SIDE_RECT;
        if (geom == GeomDisc) {
            ra.gc[0] = glob[2];
        } else {
            ra.gc[0] = 0.5 * (glob[0] + glob[2]);
            ra.gc[1] = 0.5 * (glob[1] + glob[3]);
            ra.gc[2] = glob[2] - ra.gc[0];
            ra.gc[3] = glob[3] - ra.gc[1];
        }
        const int64_t nit = (int64_t)items.size();
        for (int64_t base = 0; base < nit; base += (int64_t)1 << 30) {
            ra.item_base = base;
            hipError_t lr = launch_reduce((int)std::min<int64_t>(nit - base, (int64_t)1 << 30),
                                          kReduceWG, shmem, e->stream, ra);
            if (lr != hipSuccess) return fail(e, PDDS_ERR_DEVICE, std::string("reduce launch: ") + hipGetErrorString(lr));
        }
        FinalizeArgs f;
        f.knots = dk; f.tables = dtab; f.items = dit; f.stats = dstats; f.partials = dpart;
        f.n_knots = nk; f.S = S; f.G = dG; f.Gvar = dGv; f.kstats = dks; f.kacc = rec.kacc;
        hipError_t lr = launch_finalize(nk, 256, e->stream, f);
        if (lr != hipSuccess) return fail(e, PDDS_ERR_DEVICE, std::string("finalize launch: ") + hipGetErrorString(lr));
        HIPCHK(hipEventRecord(e->ev[2], e->stream));
        std::vector<KnotStat> ks(nk);
        HIPCHK(hipMemcpyAsync(ks.data(), dks, (size_t)nk * sizeof(KnotStat), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        for (int q = 0; q < nk; ++q) out.kstats[ord[k0 + q]] = ks[q];
        float ms1 = 0, ms2 = 0;
        HIPCHK(hipEventElapsedTime(&ms1, e->ev[0], e->ev[1]));
        HIPCHK(hipEventElapsedTime(&ms2, e->ev[1], e->ev[2]));
        kernel_ms += ms1;
        reduce_ms += ms2;
        k0 = k1;
    }
    if (n_vals > 0 && G_host) HIPCHK(hipMemcpyAsync(G_host, dG, (size_t)n_vals * 8, hipMemcpyDeviceToHost, e->stream));
    if (n_vals > 0 && Gvar_host) HIPCHK(hipMemcpyAsync(Gvar_host, dGv, (size_t)n_vals * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    out.kernel_ms = kernel_ms;
    out.reduce_ms = reduce_ms;
    out.items = items_all;
    out.grid = grid_used;
    out.P = Pmax_all;
    return PDDS_OK;
}

// Stencil tables of a mesh (device layout), built per quirk mode.
void mesh_tables(const Mesh &m, int quirk, std::vector<StencilDesc> &sten, std::vector<double> &entx,
                 std::vector<double> &enty, std::vector<int8_t> &ents, std::vector<double> &entic2,
                 std::vector<int32_t> &e1,
                 std::vector<int32_t> &e2, std::vector<double> &ipsi) {
    // one Psi^-1 block per distinct side: stencils that share a side point at the same block
    std::vector<int64_t> side_off;
    for (const Stencil &s : m.stencils) {
        StencilDesc d;
        std::memset(&d, 0, sizeof(d));
        for (int i = 0; i < 6; ++i) d.geo[i] = s.geo[i];
        for (int i = 0; i < 4; ++i) {
            d.rect[i] = s.rect[i];
            d.bbox[i] = s.bbox[i];
            d.inv_c2[i] = 1.0 / (s.cshape[i] * s.cshape[i]);
            d.msz[i] = (int)s.idx[i].size();
        }
        d.ent_off = (int32_t)entx.size();
        d.csr_nnz = (int32_t)s.cols.size();
        d.csr_off = (int32_t)e1.size();
        d.gmask = s.gmask;
        for (int i = 0; i < 4; ++i) {
            const int sid = s.side_id[i];
            if (sid >= 0 && sid < (int)side_off.size() && side_off[sid] >= 0) {
                d.ipsi_off[i] = side_off[sid];
            } else {
                d.ipsi_off[i] = (int64_t)ipsi.size();
                ipsi.insert(ipsi.end(), s.ipsi[i].begin(), s.ipsi[i].end());
                if (sid >= 0) {
                    if (sid >= (int)side_off.size()) side_off.resize(sid + 1, -1);
                    side_off[sid] = d.ipsi_off[i];
                }
            }
            for (size_t j = 0; j < s.idx[i].size(); ++j) {
                entx.push_back(s.px[i][j]);
                enty.push_back(s.py[i][j]);
                ents.push_back((int8_t)i);
                const double c = s.cent[i].empty() ? s.cshape[i] : s.cent[i][j];
                entic2.push_back(1.0 / (c * c));
            }
        }
        const auto &a1 = quirk ? s.e1 : s.e1_sum;
        for (size_t c = 0; c < s.cols.size(); ++c) {
            e1.push_back(a1[c]);
            e2.push_back(quirk ? -1 : s.e2_sum[c]);
        }
        sten.push_back(d);
    }
}

// Device-resident stencil tables of (mesh, quirk): built on first use, then reused.
// Disc meshes always sum a corner entry's sides: the "=+" quirk is a property of the
// reference's rectangular stencils, and on a disc the apex of every centre sector is such a
// corner, reached by most of the paths of the knots next to it.
pdds_status mesh_dev_tables(pdds_engine *e, const pdds_mesh *m, int quirk, DevTables **out) {
    const int q = (quirk && m->m.geom == GeomRect) ? 1 : 0;
    for (DevTables &t : e->tables)
        if (t.mesh_id == m->id && t.quirk == q) { *out = &t; return PDDS_OK; }
    std::vector<StencilDesc> sten;
    std::vector<double> entx, enty, ipsi, entic2;
    std::vector<int8_t> ents;
    std::vector<int32_t> e1, e2;
    mesh_tables(m->m, q, sten, entx, enty, ents, entic2, e1, e2, ipsi);
    DevTables &t = e->tables[e->tables_next];
    e->tables_next = (e->tables_next + 1) % 2;
    pdds_status st = build_tables(e, sten, entx, enty, ents, entic2, e1, e2, ipsi, t);
    if (st) return st;
    t.mesh_id = m->id;
    t.quirk = q;
    t.geom = m->m.geom;
    *out = &t;
    return PDDS_OK;
}

void fill_counters(pdds_counters *ctr, const AsmOut &o, double total_ms) {
    if (!ctr) return;
    std::memset(ctr, 0, sizeof(*ctr));
    for (const KnotStat &k : o.kstats) {
        ctr->path_steps += (int64_t)k.steps;
        ctr->capped += (int64_t)k.capped;
        for (int q = 0; q < 6; ++q) ctr->exits[q] += (int64_t)k.cnt[q];
    }
    for (int q = 0; q < 6; ++q) ctr->paths += ctr->exits[q];
    ctr->kernel_ms = o.kernel_ms;
    ctr->reduce_ms = o.reduce_ms;
    ctr->total_ms = total_ms;
    ctr->work_items = o.items;
    ctr->grid_wgs = o.grid;
    ctr->wg_size = o.wg;
    ctr->item_paths = o.P;
    ctr->sub_batch = o.S;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" {

const char *pdds_status_string(pdds_status s) {
    switch (s) {
        case PDDS_OK: return "ok";
        case PDDS_ERR_INVALID: return "invalid argument";
        case PDDS_ERR_NOMEM: return "out of memory";
        case PDDS_ERR_DEVICE: return "device error";
        case PDDS_ERR_UNSUPPORTED: return "unsupported";
        case PDDS_ERR_SINGULAR: return "singular interpolation matrix";
        case PDDS_ERR_NOT_READY: return "not ready";
        case PDDS_ERR_FULL: return "queue full or over budget";
        default: return "unknown status";
    }
}

void pdds_run_defaults(pdds_run *r) {
    if (!r) return;
    std::memset(r, 0, sizeof(*r));
    r->h = 1e-3;
    r->beta = -0.5826;
    r->seed = 42;
    r->precision = PDDS_FP64;
    r->quirk_last_side = 1;
    r->lut_lds = 1;
}

pdds_status pdds_engine_create(int device, pdds_engine **out) {
    if (!out) return PDDS_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PDDS_ERR_DEVICE;
    if (device < 0 || device >= n) return PDDS_ERR_INVALID;
    pdds_engine *e = new pdds_engine();
    e->device = device;
    e->bufs.resize(B_NSLOTS);
    // on a failure part way, pdds_engine_destroy releases what was created (members start null)
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        e->stream = nullptr;
        pdds_engine_destroy(e);
        return PDDS_ERR_DEVICE;
    }
    for (auto &ev : e->ev)
        if (hipEventCreate(&ev) != hipSuccess) {
            ev = nullptr;
            pdds_engine_destroy(e);
            return PDDS_ERR_DEVICE;
        }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        e->n_cu = prop.multiProcessorCount;
        e->name = prop.gcnArchName;
        e->lds_max = std::max<size_t>({(size_t)prop.sharedMemPerBlock, (size_t)prop.maxSharedMemoryPerMultiProcessor,
                                        (size_t)64 * 1024});
        e->lds_max = std::min<size_t>(e->lds_max, (size_t)160 * 1024);
    }
    if (e->n_cu <= 0) e->n_cu = 256;
    // this device's no-allocation scratch array (looked up once per device); without it only
    // pdds_point_eval_noalloc is unavailable
    char *sp = nullptr;
    if (noalloc_scratch_of(device, &sp) == hipSuccess) e->scratch = sp;
    *out = e;
    return PDDS_OK;
}

void pdds_engine_destroy(pdds_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &b : e->bufs)
        if (b.p) (void)hipFree(b.p);
    for (auto &t : e->tables) t.release();
    e->dom.release();
    for (auto &ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *pdds_last_error(const pdds_engine *e) { return e ? e->err.c_str() : "null engine"; }

pdds_status pdds_engine_info(const pdds_engine *e, int32_t *n_cu, char *name, int32_t name_len) {
    if (!e) return PDDS_ERR_INVALID;
    if (n_cu) *n_cu = e->n_cu;
    if (name && name_len > 0) {
        std::strncpy(name, e->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return PDDS_OK;
}

// ------------------------------------------------------------------ LUT ----
pdds_status pdds_lut_create(pdds_engine *e, int32_t nx, int32_t ny, int32_t nf, const double *samples, double x0,
                            double y0, double x1, double y1, int32_t mode, pdds_lut **out) {
    if (!e || !out || !samples) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    *out = nullptr;
    if (nf < 1 || nf > 8 || nx < 4 || ny < 4 || !(x1 > x0) || !(y1 > y0) || (mode != 0 && mode != 1))
        return fail(e, PDDS_ERR_INVALID, "invalid LUT shape");
    std::vector<float> coef;
    LutView v{};
    int NX, NY;
    if (mode == 0) {
        NX = nx + 2; NY = ny + 2;
        coef.resize((size_t)NX * NY * nf);
        lut_prefilter(nx, ny, nf, samples, coef.data());
        double dx = (x1 - x0) / (double)(nx - 1), dy = (y1 - y0) / (double)(ny - 1);
        v.ox = x0 - dx; v.oy = y0 - dy; v.sx = 1.0 / dx; v.sy = 1.0 / dy;
    } else {
        NX = nx; NY = ny;
        coef.resize((size_t)NX * NY * nf);
        for (size_t q = 0; q < coef.size(); ++q) coef[q] = (float)samples[q];
        v.sx = (double)nx / (x1 - x0); v.sy = (double)ny / (y1 - y0);
        v.ox = x0 + 0.5 / v.sx; v.oy = y0 + 0.5 / v.sy;
    }
    // device layout (see LutView): replicated border of kLutPad texels below and
    // kLutPad + 1 above in both directions; fields 0..3 in a float4 plane; field 4 alone
    // (nf == 5, a float plane) or fields 4..7 per texel (nf > 5, a float4 plane) in plane b
    const int PX = NX + 2 * kLutPad + 1, PY = NY + 2 * kLutPad + 1;
    const size_t nt = (size_t)PX * PY;
    const bool has_b = nf > 4;
    if (nt * 16 >= ((size_t)1 << 32)) return fail(e, PDDS_ERR_UNSUPPORTED, "LUT larger than 4 GiB per plane");
    std::vector<float> pa(nt * 4, 0.0f), pb(nf > 5 ? nt * 4 : nf == 5 ? nt : 4, 0.0f);
    auto src = [&](int pj, int pi, int f) {
        const int j = std::min(std::max(pj - kLutPad, 0), NY - 1);
        const int i = std::min(std::max(pi - kLutPad, 0), NX - 1);
        return coef[((size_t)j * NX + i) * nf + f];
    };
    for (int pj = 0; pj < PY; ++pj)
        for (int pi = 0; pi < PX; ++pi) {
            const size_t t = (size_t)pj * PX + pi;
            for (int f = 0; f < std::min(nf, 4); ++f) pa[t * 4 + f] = src(pj, pi, f);
            if (nf == 5)
                pb[t] = src(pj, pi, 4);
            else
                for (int f = 4; f < nf; ++f) pb[t * 4 + (f - 4)] = src(pj, pi, f);
        }
    v.nx = NX; v.ny = NY; v.px = PX; v.nf = nf;
    pdds_lut *l = new pdds_lut();
    l->e = e;
    const size_t bytes_a = pa.size() * sizeof(float), bytes_b = has_b ? pb.size() * sizeof(float) : 0;
    l->host.resize((bytes_a + bytes_b) / sizeof(float));
    std::memcpy(l->host.data(), pa.data(), bytes_a);
    if (bytes_b) std::memcpy(l->host.data() + pa.size(), pb.data(), bytes_b);
    l->hview = v;
    l->hview.a = (const float4 *)l->host.data();
    l->hview.b = bytes_b ? (const void *)(l->host.data() + pa.size()) : nullptr;
    if (hipSetDevice(e->device) != hipSuccess || hipMalloc(&l->d, bytes_a + bytes_b) != hipSuccess) {
        delete l;
        return fail(e, PDDS_ERR_NOMEM, "LUT allocation failed");
    }
    if (hipMemcpy(l->d, pa.data(), bytes_a, hipMemcpyHostToDevice) != hipSuccess ||
        (bytes_b && hipMemcpy((char *)l->d + bytes_a, pb.data(), bytes_b, hipMemcpyHostToDevice) != hipSuccess)) {
        (void)hipFree(l->d);
        delete l;
        return fail(e, PDDS_ERR_DEVICE, "LUT upload failed");
    }
    v.a = (const float4 *)l->d;
    v.b = bytes_b ? (const void *)((const char *)l->d + bytes_a) : nullptr;
    l->view = v;
    l->nx = nx; l->ny = ny; l->nf = nf; l->mode = mode;
    *out = l;
    return PDDS_OK;
}

void pdds_lut_destroy(pdds_lut *l) {
    if (!l) return;
    if (l->d) (void)hipFree(l->d);
    delete l;
}

pdds_status pdds_lut_eval(pdds_engine *e, const pdds_lut *l, int32_t field, int64_t n, const double *x,
                          const double *y, double *out) {
    if (!e || !l || !x || !y || !out || n < 0 || field < 0 || field >= l->nf) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (n == 0) return PDDS_OK;
    void *px, *py, *po;
    pdds_status st;
    if ((st = dbuf(e, B_P0, (size_t)n * 8, &px)) || (st = dbuf(e, B_P1, (size_t)n * 8, &py)) ||
        (st = dbuf(e, B_P2, (size_t)n * 8, &po)))
        return st;
    HIPCHK(hipMemcpyAsync(px, x, (size_t)n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(py, y, (size_t)n * 8, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_lut_eval(l->view, field, n, (const double *)px, (const double *)py, (double *)po, e->stream));
    HIPCHK(hipMemcpyAsync(out, po, (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return PDDS_OK;
}

// ----------------------------------------------------------------- mesh ----
pdds_status pdds_mesh_create(double sw_x, double sw_y, double ne_x, double ne_y, int32_t n_sub, int32_t n_node,
                             const double fac[4], pdds_mesh **out) {
    if (!out || !fac) return PDDS_ERR_INVALID;
    *out = nullptr;
    if (n_sub < 2 || n_node < 4 || n_sub > 4096 || n_node > 100000 || !(ne_x > sw_x) || !(ne_y > sw_y))
        return PDDS_ERR_INVALID;
    for (int i = 0; i < 4; ++i)
        if (!(fac[i] > 0.0)) return PDDS_ERR_INVALID;
    pdds_mesh *m = new pdds_mesh(sw_x, sw_y, ne_x, ne_y, n_sub, n_node, fac);
    *out = m;
    return m->m.any_singular ? PDDS_ERR_SINGULAR : PDDS_OK;
}

pdds_status pdds_mesh_create_disc(double radius, int32_t n_rings, int32_t n_sectors, uint32_t seed, int32_t kmin,
                                  int32_t kmax, double shape, pdds_mesh **out) {
    if (!out) return PDDS_ERR_INVALID;
    *out = nullptr;
    if (!(radius > 0.0) || n_rings < 2 || n_rings > 1024 || n_sectors < 6 || n_sectors > 4096 || kmin < 1 ||
        kmax < kmin || kmax > 100000 || !(shape > 0.0))
        return PDDS_ERR_INVALID;
    pdds_mesh *m = new pdds_mesh(Mesh::disc(radius, n_rings, n_sectors, seed, kmin, kmax, shape));
    *out = m;
    return m->m.any_singular ? PDDS_ERR_SINGULAR : PDDS_OK;
}

pdds_status pdds_mesh_geometry(const pdds_mesh *m, int32_t knot, int32_t *geom, double domain[6]) {
    if (!m || knot < 0 || knot >= m->m.n_knots()) return PDDS_ERR_INVALID;
    const Mesh &M = m->m;
    if (geom) *geom = M.geom;
    if (domain) {
        for (int i = 0; i < 6; ++i) domain[i] = 0.0;
        const int t = M.stencil_of[knot];
        if (t >= 0) {
            const Stencil &s = M.stencils[t];
            if (M.geom == GeomDisc) for (int i = 0; i < 6; ++i) domain[i] = s.geo[i];
            else for (int i = 0; i < 4; ++i) domain[i] = s.rect[i];
        }
    }
    return PDDS_OK;
}

pdds_status pdds_mesh_entry_shapes(const pdds_mesh *m, int32_t knot, int32_t cap, double *c, int32_t *n_entries) {
    if (!m || knot < 0 || knot >= m->m.n_knots()) return PDDS_ERR_INVALID;
    const int t = m->m.stencil_of[knot];
    int tot = t >= 0 ? m->m.stencils[t].total() : 0;
    if (n_entries) *n_entries = tot;
    if (c) {
        if (cap < tot) return PDDS_ERR_INVALID;
        int o = 0;
        for (int d = 0; t >= 0 && d < 4; ++d) {
            const Stencil &s = m->m.stencils[t];
            for (size_t j = 0; j < s.idx[d].size(); ++j) c[o++] = s.cent[d].empty() ? s.cshape[d] : s.cent[d][j];
        }
    }
    return PDDS_OK;
}

pdds_status pdds_mesh_shape(const pdds_mesh *m, int32_t knot, double c[4]) {
    if (!m || !c || knot < 0 || knot >= m->m.n_knots()) return PDDS_ERR_INVALID;
    const int t = m->m.stencil_of[knot];
    for (int i = 0; i < 4; ++i) c[i] = t >= 0 ? m->m.stencils[t].cshape[i] : 0.0;
    return PDDS_OK;
}

void pdds_mesh_destroy(pdds_mesh *m) { delete m; }
int32_t pdds_mesh_n_knots(const pdds_mesh *m) { return m ? m->m.n_knots() : 0; }
int32_t pdds_mesh_n_stencils(const pdds_mesh *m) { return m ? (int32_t)m->m.stencils.size() : 0; }

pdds_status pdds_mesh_knots(const pdds_mesh *m, double *x0, double *y0, double *px, double *py,
                            int32_t *on_boundary, int32_t *stencil_id) {
    if (!m) return PDDS_ERR_INVALID;
    const Mesh &M = m->m;
    for (int q = 0; q < M.n_knots(); ++q) {
        if (x0) x0[q] = M.x0[q];
        if (y0) y0[q] = M.y0[q];
        if (px) px[q] = M.npx[q];
        if (py) py[q] = M.npy[q];
        if (on_boundary) on_boundary[q] = M.on_boundary[q];
        if (stencil_id) stencil_id[q] = M.stencil_of[q];
    }
    return PDDS_OK;
}

pdds_status pdds_mesh_stencil(const pdds_mesh *m, int32_t knot, double rect[4], int32_t msz[4], int32_t kind[4],
                              int32_t cap, int32_t *idx, double *px, double *py, int32_t *n_entries) {
    if (!m || knot < 0 || knot >= m->m.n_knots()) return PDDS_ERR_INVALID;
    static const Stencil none{};
    const Stencil &s = m->m.stencil_of[knot] >= 0 ? m->m.stencils[m->m.stencil_of[knot]] : none;
    int tot = s.total();
    if (n_entries) *n_entries = tot;
    for (int i = 0; i < 4; ++i) {
        if (rect) rect[i] = s.rect[i];
        if (msz) msz[i] = (int32_t)s.idx[i].size();
        if (kind) kind[i] = s.kind[i];
    }
    if (idx || px || py) {
        if (cap < tot) return PDDS_ERR_INVALID;
        int o = 0;
        for (int d = 0; d < 4; ++d)
            for (size_t j = 0; j < s.idx[d].size(); ++j, ++o) {
                if (idx) idx[o] = s.idx[d][j];
                if (px) px[o] = s.px[d][j];
                if (py) py[o] = s.py[d][j];
            }
    }
    return PDDS_OK;
}

pdds_status pdds_mesh_cond(const pdds_mesh *m, int32_t knot, double cond[4]) {
    if (!m || !cond || knot < 0 || knot >= m->m.n_knots()) return PDDS_ERR_INVALID;
    static const Stencil none{};
    const Stencil &s = m->m.stencil_of[knot] >= 0 ? m->m.stencils[m->m.stencil_of[knot]] : none;
    for (int i = 0; i < 4; ++i) cond[i] = s.cond[i];
    return PDDS_OK;
}

pdds_status pdds_mesh_ipsi(const pdds_mesh *m, int32_t knot, int32_t side, int64_t cap, double *out, int32_t *msz) {
    if (!m || knot < 0 || knot >= m->m.n_knots() || side < 0 || side > 3) return PDDS_ERR_INVALID;
    static const Stencil none{};
    const Stencil &s = m->m.stencil_of[knot] >= 0 ? m->m.stencils[m->m.stencil_of[knot]] : none;
    const int64_t mm = (int64_t)s.idx[side].size();
    if (msz) *msz = (int32_t)mm;
    if (out) {
        if (cap < mm * mm) return PDDS_ERR_INVALID;
        std::copy(s.ipsi[side].begin(), s.ipsi[side].end(), out);
    }
    return PDDS_OK;
}

pdds_status pdds_mesh_csr(const pdds_mesh *m, int64_t *row_ptr, int32_t *cols, int64_t *nnz) {
    if (!m) return PDDS_ERR_INVALID;
    const Mesh &M = m->m;
    if (nnz) *nnz = M.row_ptr.back();
    if (row_ptr) std::copy(M.row_ptr.begin(), M.row_ptr.end(), row_ptr);
    if (cols) std::copy(M.cols.begin(), M.cols.end(), cols);
    return PDDS_OK;
}

// ------------------------------------------------------------- assembly ----
// Every Monte Carlo knot of mesh m in one batch (the worker loop of PDDSparseGM::Solve,
// PDDAlgorithms/PDDSparseGM.cpp:795-848); with u_ref also each knot's equation estimator
// B_k - sum_j G_kj u_ref_j (mean and per-path variance).  Caller holds e->mu.
static pdds_status assemble_locked(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                                   const int64_t *paths_per_knot, int64_t n_paths, const double *u_ref, double *G,
                                   double *G_var, double *B, double *BB, double *eq_mean, double *eq_var,
                                   AsmOut &o) {
    if (!m || !B || !BB) return fail(e, PDDS_ERR_INVALID, "null mesh or output");
    pdds_status st = check_run(e, r);
    if (st) return st;
    const Mesh &M = m->m;
    if (M.any_singular) return fail(e, PDDS_ERR_SINGULAR, "mesh has a singular Psi");
    PdeParams pp;
    if ((st = make_params(e, p, r, pp))) return st;
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    DevTables *T;
    if ((st = mesh_dev_tables(e, m, r->quirk_last_side, &T))) return st;
    std::vector<AsmKnot> knots;
    std::vector<int> kmap, skipped;
    for (int q = 0; q < M.n_knots(); ++q) {
        int64_t n = paths_per_knot ? paths_per_knot[q] : n_paths;
        if (n < 0) return fail(e, PDDS_ERR_INVALID, "negative path count");
        if (M.on_boundary[q] && pp.refl) {
            // a knot on a reflecting side has no Dirichlet value (it would need a stencil of its own)
            const double tol = 1e-12 * (1.0 + std::fabs(M.glob[2] - M.glob[0]));
            const bool on[4] = {std::fabs(M.y0[q] - M.glob[3]) < tol, std::fabs(M.y0[q] - M.glob[1]) < tol,
                                std::fabs(M.x0[q] - M.glob[2]) < tol, std::fabs(M.x0[q] - M.glob[0]) < tol};
            for (int s = 0; s < 4; ++s)
                if (on[s] && ((pp.refl >> s) & 1))
                    return fail(e, PDDS_ERR_UNSUPPORTED, "a knot lies on a reflecting side (knot " + std::to_string(q) + ")");
        }
        if (M.on_boundary[q]) {
            B[q] = host_g(p, M.x0[q], M.y0[q], pp);
            BB[q] = B[q] * B[q];
            if (eq_mean) eq_mean[q] = B[q];
            if (eq_var) eq_var[q] = 0.0;
            continue;
        }
        if (n == 0) {
            // with per-knot counts, 0 leaves the knot to another shard of a split by knots
            if (!paths_per_knot) return fail(e, PDDS_ERR_INVALID, "n_paths must be positive");
            skipped.push_back(q);
            continue;
        }
        AsmKnot k{M.x0[q], M.y0[q], (uint32_t)q, M.stencil_of[q], n, M.row_ptr[q]};
        knots.push_back(k);
        kmap.push_back(q);
    }
    std::vector<double> W;
    if (u_ref) {
        W.resize(M.cols.size());
        for (size_t i = 0; i < M.cols.size(); ++i) W[i] = u_ref[M.cols[i]];
    }
    st = run_assembly(e, p, r, M.glob, {T}, knots, M.row_ptr.back(), G, G_var, o, u_ref ? W.data() : nullptr);
    if (st) return st;
    for (size_t i = 0; i < knots.size(); ++i) {
        B[kmap[i]] = o.kstats[i].B;
        BB[kmap[i]] = o.kstats[i].BB;
        if (eq_mean) eq_mean[kmap[i]] = u_ref ? o.kstats[i].Rm : 0.0;
        if (eq_var) eq_var[kmap[i]] = u_ref ? o.kstats[i].Rv : 0.0;
    }
    for (int q : skipped) {
        for (int64_t i = M.row_ptr[q]; i < M.row_ptr[q + 1]; ++i) {
            if (G) G[i] = 0.0;
            if (G_var) G_var[i] = 0.0;
        }
        B[q] = 0.0;
        BB[q] = 0.0;
        if (eq_mean) eq_mean[q] = 0.0;
        if (eq_var) eq_var[q] = 0.0;
    }
    return PDDS_OK;
}

pdds_status pdds_assemble(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                          const int64_t *paths_per_knot, int64_t n_paths, double *G, double *G_var, double *B,
                          double *BB, pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const double t0 = now_ms();
    AsmOut o;
    pdds_status st = assemble_locked(e, m, p, r, paths_per_knot, n_paths, nullptr, G, G_var, B, BB, nullptr,
                                     nullptr, o);
    if (st) return st;
    fill_counters(ctr, o, now_ms() - t0);
    return PDDS_OK;
}

pdds_status pdds_assemble_eq(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                             const int64_t *paths_per_knot, int64_t n_paths, const double *u_ref, double *G,
                             double *G_var, double *B, double *BB, double *eq_mean, double *eq_var,
                             pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const double t0 = now_ms();
    if (!u_ref || !eq_var) return fail(e, PDDS_ERR_INVALID, "u_ref and eq_var are required");
    AsmOut o;
    pdds_status st = assemble_locked(e, m, p, r, paths_per_knot, n_paths, u_ref, G, G_var, B, BB, eq_mean, eq_var, o);
    if (st) return st;
    fill_counters(ctr, o, now_ms() - t0);
    return PDDS_OK;
}

// ------------------------------------------------- variance-targeted path counts ----
void pdds_adaptive_defaults(pdds_adaptive *c) {
    if (!c) return;
    c->se_target = 1e-3;
    c->n_min = 1000;
    c->n_max = 1000000;
    c->n_pilot = 0;
    c->max_rounds = 3;
    c->tol = 1e-10;
    c->growth = 1.1;
    c->variance = 0;
}

// Pilot assembly, then rounds of: assemble with the current counts, solve, and move each
// knot's count towards N_k = growth * var_k / se^2 (within [n_min, n_max]), var_k the
// batch-means variance of its score (variance = 0) or of its equation weighed with a reference
// solution (variance = 1), sqrt(var_k / N_k) being its standard error.  Each round draws fresh
// paths (a Philox stream of its own), so a round's counts do not depend on its own paths, and
// the final G, B and u are those of one plain assembly with the returned counts and stream.
// The policy stands in for the commented Compute_h_N of the reference
// (PDDAlgorithms/PDDSparseGM.cpp:4338-4395, N_i = 4 Var_i (1 - rho_i^2) / (0.6667 eps)^2, the
// variance of the control-variate estimator with a standard error of eps/3): variance = 0 is the
// batch-means variance of our score, whose control variate is built in.
pdds_status pdds_solve_adaptive(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                                const pdds_adaptive *cfg_in, double *u, int64_t *paths_per_knot, double *se,
                                pdds_adaptive_report *rep) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const double t0 = now_ms();
    if (!m || !r || !u || !paths_per_knot) return fail(e, PDDS_ERR_INVALID, "null mesh, run or output");
    pdds_adaptive cfg;
    pdds_adaptive_defaults(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (!(cfg.se_target > 0) || cfg.n_min < 1 || cfg.n_max < cfg.n_min || cfg.max_rounds < 1 || cfg.max_rounds > 64 ||
        !(cfg.growth >= 1.0) || cfg.variance < 0 || cfg.variance > 1)
        return fail(e, PDDS_ERR_INVALID, "bad adaptive configuration");
    const int64_t n_pilot = cfg.n_pilot > 0 ? cfg.n_pilot : cfg.n_min;
    const Mesh &M = m->m;
    const int n = M.n_knots();
    const int64_t nnz = M.row_ptr.back();
    std::vector<double> G(std::max<int64_t>(nnz, 1)), Gv(std::max<int64_t>(nnz, 1)), B(n), BB(n), em(n), ev(n),
        uref(n);
    std::vector<int64_t> N(n, n_pilot);
    pdds_adaptive_report R;
    std::memset(&R, 0, sizeof(R));
    const double tol = cfg.tol > 0 ? cfg.tol : 1e-10;
    int32_t it = 0;
    double rr = 0.0;
    // every round draws fresh paths: round a uses stream r->stream + a * kRoundStream
    constexpr uint32_t kRoundStream = 0x9E3779B9u;
    pdds_run run = *r;
    auto assemble = [&](const double *ur) -> pdds_status {
        AsmOut o;
        pdds_status st = assemble_locked(e, m, p, &run, N.data(), 0, ur, G.data(), Gv.data(), B.data(), BB.data(),
                                         em.data(), ev.data(), o);
        if (st) return st;
        for (const KnotStat &k : o.kstats) R.path_steps += (int64_t)k.steps;
        R.kernel_ms += o.kernel_ms + o.reduce_ms;
        return PDDS_OK;
    };
    auto solve = [&](double *out) -> pdds_status {
        return solve_system(e, n, M.row_ptr.data(), M.cols.data(), G.data(), B.data(), out, tol, 0, &it, &rr);
    };
    // The equations are weighed with a reference solution.  A noisy reference inflates the
    // measured variance: the RBF cardinal functions of the stencil sides amplify its knot-to-knot
    // noise at the exit points.  Every round draws fresh paths, so the rounds' solutions are
    // independent estimates of u, and the reference is their mean weighted by each round's paths,
    // whose noise falls as rounds accumulate.  Counts move towards growth * var / se^2 by at most
    // a factor 8 per round, up or down, so that an over-estimate from an early, noisy reference is
    // taken back once the reference improves.
    // With variance = 0 the weights are zero, so that the equation statistics are the score's.
    const bool eq_var = cfg.variance == 1;
    pdds_status st;
    if (eq_var) {
        if ((st = assemble(nullptr)) || (st = solve(uref.data())) || (st = assemble(uref.data()))) return st;
    } else {
        std::fill(uref.begin(), uref.end(), 0.0);
        if ((st = assemble(uref.data()))) return st;
    }
    std::vector<double> usum(n, 0.0);
    double wsum = 0.0;
    auto add_solution = [&](const double *v) {
        double w = 0.0;
        for (int q = 0; q < n; ++q)
            if (!M.on_boundary[q]) w += (double)N[q];
        for (int q = 0; q < n; ++q) usum[q] += w * v[q];
        wsum += w;
    };
    if (eq_var) add_solution(uref.data());
    auto target = [&](int q) -> int64_t {
        const double want = std::ceil(cfg.growth * ev[q] / (cfg.se_target * cfg.se_target));
        double c = std::min(std::max(want, (double)cfg.n_min), (double)cfg.n_max);
        c = std::min(c, 8.0 * (double)N[q]);
        c = std::max(c, std::max((double)cfg.n_min, std::floor((double)N[q] / 8.0)));
        return (int64_t)c;
    };
    for (int q = 0; q < n; ++q)
        if (!M.on_boundary[q]) N[q] = target(q);
    for (int round = 0; round < cfg.max_rounds; ++round) {
        run.stream = r->stream + (uint32_t)(round + 1) * kRoundStream;
        if ((st = assemble(uref.data())) || (st = solve(u))) return st;
        if (eq_var) add_solution(u);
        R.rounds = round + 1;
        bool over = false, moved = false;
        std::vector<int64_t> next(N);
        for (int q = 0; q < n; ++q) {
            if (M.on_boundary[q]) continue;
            const double sek = std::sqrt(std::max(ev[q], 0.0) / (double)N[q]);
            if (sek > cfg.se_target && N[q] < cfg.n_max) over = true;
            next[q] = target(q);
            // a knot within target keeps its count unless the target is well below it
            if (sek <= cfg.se_target && next[q] > N[q]) next[q] = N[q];
            if (sek <= cfg.se_target && next[q] * 2 > N[q]) next[q] = N[q];
            moved = moved || next[q] != N[q];
        }
        // the counts returned are those of the last assembly
        if ((!over && !moved) || round + 1 == cfg.max_rounds) break;
        if (eq_var && wsum > 0.0)
            for (int q = 0; q < n; ++q) uref[q] = usum[q] / wsum;
        N.swap(next);
    }
    R.paths = 0;
    R.max_se = 0.0;
    R.knots_capped = 0;
    R.knots_over = 0;
    for (int q = 0; q < n; ++q) {
        paths_per_knot[q] = M.on_boundary[q] ? 0 : N[q];
        const double sek = M.on_boundary[q] ? 0.0 : std::sqrt(std::max(ev[q], 0.0) / (double)N[q]);
        if (se) se[q] = sek;
        if (M.on_boundary[q]) continue;
        R.paths += N[q];
        R.max_se = std::max(R.max_se, sek);
        if (sek > cfg.se_target) R.knots_over++;
        if (sek > cfg.se_target && N[q] >= cfg.n_max) R.knots_capped++;
    }
    R.iters = it;
    R.relres = rr;
    R.stream = run.stream;
    R.total_ms = now_ms() - t0;
    if (rep) *rep = R;
    return PDDS_OK;
}

// ------------------------------------------------------------- autotuner ----
pdds_status pdds_autotune(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *base,
                          const pdds_tune_config *cfg_in, pdds_run *best, pdds_tune_report *rep) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const double t0 = now_ms();
    if (!m || !p || !best) return fail(e, PDDS_ERR_INVALID, "null mesh, problem or output");
    pdds_status st = check_run(e, base);
    if (st) return st;
    pdds_tune_config cfg{16384, 512, 2};
    if (cfg_in) {
        if (cfg_in->sample_paths > 0) cfg.sample_paths = cfg_in->sample_paths;
        if (cfg_in->max_knots > 0) cfg.max_knots = cfg_in->max_knots;
        if (cfg_in->reps > 0) cfg.reps = cfg_in->reps;
    }
    const Mesh &M = m->m;
    if (M.any_singular) return fail(e, PDDS_ERR_SINGULAR, "mesh has a singular Psi");
    PdeParams pp;
    if ((st = make_params(e, p, base, pp))) return st;
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    DevTables *T;
    if ((st = mesh_dev_tables(e, m, base->quirk_last_side, &T))) return st;
    std::vector<int> mc;
    for (int q = 0; q < M.n_knots(); ++q)
        if (!M.on_boundary[q]) mc.push_back(q);
    if (mc.empty()) return fail(e, PDDS_ERR_INVALID, "mesh has no Monte Carlo knot");
    // a contiguous run of knots from the middle of the mesh: neighbouring knots share stencils
    // and LUT tiles, so the sample has the patch structure of the whole mesh
    const size_t nk = std::min(mc.size(), (size_t)cfg.max_knots);
    const size_t k0 = (mc.size() - nk) / 2;
    std::vector<AsmKnot> knots;
    for (size_t i = k0; i < k0 + nk; ++i) {
        const int q = mc[i];
        knots.push_back(AsmKnot{M.x0[q], M.y0[q], (uint32_t)q, M.stencil_of[q], cfg.sample_paths, M.row_ptr[q]});
    }
    // candidate launch shapes (none of them changes a result bit)
    std::vector<pdds_run> cands;
    auto add = [&](int lds, int wg, int grid, int ppl, int pad) {
        pdds_run r = *base;
        r.plain = 0;
        r.lut_lds = lds;
        r.wg_size = wg;
        r.grid_wgs = grid;
        r.paths_per_lane = ppl;
        r.lut_pad = pad;
        cands.push_back(r);
    };
    {
        pdds_run r = *base;
        r.plain = 0;
        cands.push_back(r);
    }
    // LDS-tiled shapes: work-groups per CU, paths per lane and the tile's row pitch
    const bool can_tile = p->pde == PDDS_PDE_VARCOEF_LUT && pp.lut.nf == 5;
    if (can_tile)
        for (int pad : {0, 1, 2, 3})
            for (int g : {1, 2})
                for (int ppl : {1, 2, 4, 8}) add(1, 0, g * e->n_cu, ppl, pad);
    for (int wg : {64, 128, 256})
        for (int g : {1, 2, 4, 8})
            for (int ppl : {1, 4, 16}) add(0, wg, g * e->n_cu * 256 / wg, ppl, 0);
    double best_rate = -1.0, base_rate = 0.0;
    size_t best_i = 0;
    for (size_t i = 0; i < cands.size(); ++i) {
        double tmin = 1e300;
        int64_t steps = 0;
        for (int rep_i = 0; rep_i <= cfg.reps; ++rep_i) {
            AsmOut o;
            const double ts = now_ms();
            if ((st = run_assembly(e, p, &cands[i], M.glob, {T}, knots, M.row_ptr.back(), nullptr, nullptr, o)))
                return st;
            const double dt = now_ms() - ts;
            if (rep_i == 0) continue;   // warm-up
            tmin = std::min(tmin, dt);
            steps = 0;
            for (const KnotStat &k : o.kstats) steps += (int64_t)k.steps;
        }
        const double rate = (double)steps / (tmin * 1e-3);
        if (i == 0) base_rate = rate;
        if (rate > best_rate) { best_rate = rate; best_i = i; }
    }
    *best = cands[best_i];
    if (rep) {
        rep->candidates = (int32_t)cands.size();
        rep->best_rate = best_rate;
        rep->base_rate = base_rate;
        rep->total_ms = now_ms() - t0;
    }
    return PDDS_OK;
}

pdds_status pdds_knot_steps(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r, double *h) {
    if (!e || !m || !h) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    pdds_status st = check_run(e, r);
    if (st) return st;
    PdeParams pp;
    if ((st = make_params(e, p, r, pp))) return st;
    const Mesh &M = m->m;
    for (int q = 0; q < M.n_knots(); ++q) {
        const Stencil &S = M.stencils[M.stencil_of[q]];
        h[q] = M.on_boundary[q] ? 0.0 : knot_step(p, pp, r, M.geom, S.rect, S.geo, M.x0[q], M.y0[q]);
    }
    return PDDS_OK;
}

pdds_status pdds_solve_knot(pdds_engine *e, const pdds_mesh *m, const pdds_problem *p, const pdds_run *r,
                            int32_t knot, int64_t n_paths, double *G, double *G_var, double *B, double *BB,
                            pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    double t0 = now_ms();
    if (!m || !B || !BB || knot < 0 || knot >= m->m.n_knots()) return fail(e, PDDS_ERR_INVALID, "bad knot");
    pdds_status st = check_run(e, r);
    if (st) return st;
    const Mesh &M = m->m;
    PdeParams pp;
    if ((st = make_params(e, p, r, pp))) return st;
    if (M.on_boundary[knot]) {
        *B = host_g(p, M.x0[knot], M.y0[knot], pp);
        *BB = *B * *B;
        if (ctr) std::memset(ctr, 0, sizeof(*ctr));
        return PDDS_OK;
    }
    if (n_paths <= 0) return fail(e, PDDS_ERR_INVALID, "n_paths must be positive");
    if (M.stencils[M.stencil_of[knot]].singular) return fail(e, PDDS_ERR_SINGULAR, "singular Psi");
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    DevTables *T;
    if ((st = mesh_dev_tables(e, m, r->quirk_last_side, &T))) return st;
    std::vector<AsmKnot> knots = {AsmKnot{M.x0[knot], M.y0[knot], (uint32_t)knot, M.stencil_of[knot], n_paths, 0}};
    AsmOut o;
    st = run_assembly(e, p, r, M.glob, {T}, knots, (int64_t)M.stencils[M.stencil_of[knot]].cols.size(), G, G_var, o);
    if (st) return st;
    *B = o.kstats[0].B;
    *BB = o.kstats[0].BB;
    fill_counters(ctr, o, now_ms() - t0);
    return PDDS_OK;
}

// Feynman-Kac estimates of u at points (Solve_OMP_Analytic without a stencil,
// Solvers/EMFKACSolver.hpp:686-691): paths from each point stop on the rectangle glob.  The
// rectangle's one-stencil table is kept by the engine and rebuilt only when glob changes.
// noalloc: every device buffer, that table included, is taken from the device's scratch array
// and nothing is allocated; the caller holds the engine lock (and, with noalloc, the device's
// scratch lock).
static pdds_status point_eval_locked(pdds_engine *e, const pdds_problem *p, const pdds_run *r, const double glob[4],
                                     int32_t n, const double *x, const double *y, const int64_t *n_paths, double *u,
                                     double *u_var, pdds_counters *ctr, bool noalloc) {
    const double t0 = now_ms();
    if (!glob || !x || !y || !n_paths || !u || n < 0) return fail(e, PDDS_ERR_INVALID, "null argument");
    pdds_status st = check_run(e, r);
    if (st) return st;
    if (!(glob[2] > glob[0]) || !(glob[3] > glob[1])) return fail(e, PDDS_ERR_INVALID, "bad rectangle");
    PdeParams pp;
    if ((st = make_params(e, p, r, pp))) return st;
    std::vector<AsmKnot> knots;
    std::vector<int> kmap;
    int64_t paths = 0;
    for (int i = 0; i < n; ++i) {
        double ex, ey, nx, ny;
        int kd;
        double dist = rect_dist<double>(glob, x[i], y[i], ex, ey, nx, ny, kd);
        if (std::fabs(dist) <= 1e-8 || dist > 0.0) {
            // on (or outside) the boundary: the Dirichlet value, no Monte Carlo
            u[i] = host_g(p, ex, ey, pp);
            if (u_var) u_var[i] = 0.0;
            continue;
        }
        if (n_paths[i] <= 0) return fail(e, PDDS_ERR_INVALID, "n_paths must be positive");
        knots.push_back(AsmKnot{x[i], y[i], (uint32_t)i, 0, n_paths[i], 0});
        kmap.push_back(i);
        paths += n_paths[i];
    }
    if (noalloc) {
        // exit records, and per point its descriptor, statistics, step counters and work items (at
        // most 17 of them, P >= N / 16, each with its descriptor and statistics), with room for
        // alignment
        const int64_t need = paths * (int64_t)sizeof(ExitRec) +
                             (int64_t)knots.size() * (int64_t)(sizeof(KnotDesc) + sizeof(KnotStat) + 16 +
                                                                17 * (sizeof(ItemDesc) + sizeof(ItemStat))) +
                             ((int64_t)1 << 16);
        if (need > (int64_t)PDDS_NOALLOC_SCRATCH_BYTES)
            return fail(e, PDDS_ERR_FULL, std::to_string(paths) + " paths at " + std::to_string(knots.size()) +
                                              " points need about " + std::to_string(need >> 20) +
                                              " MiB of device buffers; the no-allocation scratch holds " +
                                              std::to_string(PDDS_NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
    }
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    StencilDesc d;
    std::memset(&d, 0, sizeof(d));
    for (int i = 0; i < 4; ++i) { d.rect[i] = glob[i]; d.bbox[i] = glob[i]; d.inv_c2[i] = 1.0; }
    d.gmask = 15;
    AsmOut o;
    if (noalloc) {
        Arena arena;
        arena.base = e->scratch;
        arena.cap = PDDS_NOALLOC_SCRATCH_BYTES;
        struct Scope {
            pdds_engine *e;
            ~Scope() { e->arena = nullptr; }
        } scope{e};
        e->arena = &arena;
        // the table lives in the scratch array: the stencil descriptor, and one zeroed piece
        // standing for its (empty) entry arrays, which no kernel reads
        void *ps, *pz;
        if ((st = dbuf(e, B_STEN, sizeof(StencilDesc), &ps)) || (st = dbuf(e, B_ENTX, 256, &pz))) return st;
        HIPCHK(hipMemcpyAsync(ps, &d, sizeof(d), hipMemcpyHostToDevice, e->stream));
        HIPCHK(launch_fill_u64((unsigned long long *)pz, 256 / 8, 0ULL, e->stream));
        DevTables T;
        T.sten = (StencilDesc *)ps;
        T.entx = T.enty = T.ipsi = T.entic2 = (double *)pz;
        T.ents = (int8_t *)pz;
        T.e1 = T.e2 = (int32_t *)pz;
        T.host_sten = {d};
        T.n_sten = 1;
        st = run_assembly(e, p, r, glob, {&T}, knots, 0, nullptr, nullptr, o);
        if (st) return st;
    } else {
        DevTables &T = e->dom;
        if (!T.sten || std::memcmp(e->dom_rect, glob, sizeof(e->dom_rect)) != 0) {
            std::vector<StencilDesc> sten = {d};
            std::vector<double> none;
            std::vector<int8_t> none8;
            std::vector<int32_t> none32;
            if ((st = build_tables(e, sten, none, none, none8, none, none32, none32, none, T))) return st;
            std::memcpy(e->dom_rect, glob, sizeof(e->dom_rect));
        }
        st = run_assembly(e, p, r, glob, {&T}, knots, 0, nullptr, nullptr, o);
        if (st) return st;
    }
    for (size_t i = 0; i < knots.size(); ++i) {
        u[kmap[i]] = o.kstats[i].B;
        if (u_var) u_var[kmap[i]] = o.kstats[i].BB - o.kstats[i].B * o.kstats[i].B;
    }
    fill_counters(ctr, o, now_ms() - t0);
    return PDDS_OK;
}

pdds_status pdds_point_eval(pdds_engine *e, const pdds_problem *p, const pdds_run *r, const double glob[4],
                            int32_t n, const double *x, const double *y, const int64_t *n_paths, double *u,
                            double *u_var, pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    return point_eval_locked(e, p, r, glob, n, x, y, n_paths, u, u_var, ctr, false);
}

pdds_status pdds_point_eval_noalloc(pdds_engine *e, const pdds_problem *p, const pdds_run *r, const double glob[4],
                                    int32_t n, const double *x, const double *y, const int64_t *n_paths, double *u,
                                    double *u_var, pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (!e->scratch) return fail(e, PDDS_ERR_DEVICE, "the no-allocation scratch array of this device is unavailable");
    // every engine on this device shares its array: calls wait for each other instead of racing
    std::lock_guard<std::mutex> lk2(g_scratch_call_mu[e->device]);
    return point_eval_locked(e, p, r, glob, n, x, y, n_paths, u, u_var, ctr, true);
}

// ---------------------------------------------------------------- paths ----
pdds_status pdds_paths(pdds_engine *e, const pdds_problem *p, const pdds_run *r, double x0, double y0,
                       const double rect[4], uint32_t knot, int64_t path0, int64_t n, int32_t phase2, double *xt,
                       double *yt, double *Y, double *Z, double *xi, int64_t *steps, double *xt2, double *yt2,
                       double *Y2, double *Z2, double *xi2, int64_t *steps2, pdds_counters *ctr) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    double t0 = now_ms();
    pdds_status st = check_run(e, r);
    if (st) return st;
    if (!rect || n < 0 || path0 < 0 || path0 + n > ((int64_t)1 << 32)) return fail(e, PDDS_ERR_INVALID, "bad path range");
    if (!(rect[2] > rect[0]) || !(rect[3] > rect[1])) return fail(e, PDDS_ERR_INVALID, "bad rectangle");
    PdeParams pp;
    if ((st = make_params(e, p, r, pp))) return st;
    if (n == 0) return PDDS_OK;
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    PathsArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x0 = x0; a.y0 = y0;
    for (int i = 0; i < 4; ++i) a.rect[i] = rect[i];
    a.knot = knot; a.seed = r->seed; a.stream = r->stream; a.phase2 = phase2 ? 1 : 0;
    a.path0 = path0; a.n = n; a.h = r->h; a.sqrth = std::sqrt(r->h); a.beta = r->beta;
    a.max_steps = r->max_steps > 0 ? r->max_steps : (int64_t)1 << 26;
    a.pp = pp;
    double *hd[10] = {xt, yt, Y, Z, xi, xt2, yt2, Y2, Z2, xi2};
    double **dd[10] = {&a.xt, &a.yt, &a.Y, &a.Z, &a.xi, &a.xt2, &a.yt2, &a.Y2, &a.Z2, &a.xi2};
    for (int i = 0; i < 10; ++i) {
        if (!hd[i] || (i >= 5 && !phase2)) continue;
        void *pp2;
        if ((st = dbuf(e, B_P0 + i, (size_t)n * 8, &pp2))) return st;
        *dd[i] = (double *)pp2;
    }
    // step counts always gathered (counters need them)
    void *ps1, *ps2;
    if ((st = dbuf(e, B_P10, (size_t)n * 8, &ps1))) return st;
    a.steps = (int64_t *)ps1;
    if (phase2) {
        if ((st = dbuf(e, B_P11, (size_t)n * 8, &ps2))) return st;
        a.steps2 = (int64_t *)ps2;
    }
    const int wg = r->wg_size > 0 ? r->wg_size : 256;
    const int64_t grid = (n + wg - 1) / wg;
    HIPCHK(hipEventRecord(e->ev[0], e->stream));
    hipError_t lr = r->precision == PDDS_FP64 ? launch_paths_f64(p->pde, (int)grid, wg, e->stream, a)
                                              : launch_paths_f32(p->pde, (int)grid, wg, e->stream, a);
    if (lr != hipSuccess) return fail(e, PDDS_ERR_DEVICE, std::string("paths launch: ") + hipGetErrorString(lr));
    HIPCHK(hipEventRecord(e->ev[1], e->stream));
    for (int i = 0; i < 10; ++i)
        if (*dd[i]) HIPCHK(hipMemcpyAsync(hd[i], *dd[i], (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    std::vector<int64_t> hs((size_t)n);
    HIPCHK(hipMemcpyAsync(hs.data(), a.steps, (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    if (phase2 && steps2) HIPCHK(hipMemcpyAsync(steps2, a.steps2, (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (steps) std::copy(hs.begin(), hs.end(), steps);
    if (ctr) {
        std::memset(ctr, 0, sizeof(*ctr));
        for (int64_t s : hs) ctr->path_steps += s;
        ctr->paths = n;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
        ctr->kernel_ms = ms;
        ctr->total_ms = now_ms() - t0;
        ctr->wg_size = wg;
        ctr->grid_wgs = (int32_t)grid;
    }
    return PDDS_OK;
}

}  // extern "C"

extern "C" pdds_status pdds_solve(pdds_engine *e, int32_t n, const int64_t *row_ptr, const int32_t *cols,
                                  const double *G, const double *B, double *u, double tol, int32_t maxit,
                                  int32_t *iters, double *relres) {
    if (!e) return PDDS_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    return solve_system(e, n, row_ptr, cols, G, B, u, tol, maxit, iters, relres);
}

// (I + G) u = B on the device; the caller holds the engine lock
pdds_status solve_system(pdds_engine *e, int32_t n, const int64_t *row_ptr, const int32_t *cols, const double *G,
                         const double *B, double *u, double tol, int32_t maxit, int32_t *iters, double *relres) {
    if (n <= 0 || !row_ptr || !B || !u || (row_ptr[n] > 0 && (!cols || !G)) || !(tol >= 0.0))
        return fail(e, PDDS_ERR_INVALID, "bad system");
    const int64_t nnz = row_ptr[n];
    for (int i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return fail(e, PDDS_ERR_INVALID, "row_ptr not monotone");
    for (int64_t k = 0; k < nnz; ++k)
        if (cols[k] < 0 || cols[k] >= n) return fail(e, PDDS_ERR_INVALID, "column out of range");
    if (hipSetDevice(e->device) != hipSuccess) return fail(e, PDDS_ERR_DEVICE, "hipSetDevice");
    void *drp, *dci, *dv, *db, *dx, *dw;
    pdds_status st;
    const int mi = maxit > 0 ? maxit : 2 * n;
    if ((st = dbuf(e, B_P0, (size_t)(n + 1) * 8, &drp)) || (st = dbuf(e, B_P1, (size_t)std::max<int64_t>(nnz, 1) * 4, &dci)) ||
        (st = dbuf(e, B_P2, (size_t)std::max<int64_t>(nnz, 1) * 8, &dv)) || (st = dbuf(e, B_P3, (size_t)n * 8, &db)) ||
        (st = dbuf(e, B_P4, (size_t)n * 8, &dx)) || (st = dbuf(e, B_P5, bicgstab_work_bytes(n, mi), &dw)))
        return st;
    HIPCHK(hipMemcpyAsync(drp, row_ptr, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, e->stream));
    if (nnz > 0) {
        HIPCHK(hipMemcpyAsync(dci, cols, (size_t)nnz * 4, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(dv, G, (size_t)nnz * 8, hipMemcpyHostToDevice, e->stream));
    }
    HIPCHK(hipMemcpyAsync(db, B, (size_t)n * 8, hipMemcpyHostToDevice, e->stream));
    struct { int it; int pad; double rr; } hr;
    HIPCHK(bicgstab_device(n, (const int64_t *)drp, (const int32_t *)dci, (const double *)dv, (const double *)db,
                           (double *)dx, dw, tol, mi, 8, &hr.it, &hr.rr, e->stream));
    HIPCHK(hipMemcpyAsync(u, dx, (size_t)n * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (iters) *iters = hr.it;
    if (relres) *relres = hr.rr;
    return PDDS_OK;
}
