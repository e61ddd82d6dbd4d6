// engine_internal.h -- shared internals of the engine library (engine.cpp, server.cpp):
// the engine object, grow-only device buffers, device stencil tables and the generic
// assembly over a list of knots.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdds_engine.h"
#include "engine_types.h"
#include "mesh.h"

using namespace pdds;

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

// Stencil tables of one mesh in device memory (kept resident across calls).
struct DevTables {
    uint64_t mesh_id = 0;
    int quirk = -1;
    int geom = 0;                         // GeomRect / GeomDisc stopping domains
    int n_sten = 0;
    StencilDesc *sten = nullptr;
    double *entx = nullptr, *enty = nullptr, *ipsi = nullptr, *entic2 = nullptr;
    int8_t *ents = nullptr;
    int32_t *e1 = nullptr, *e2 = nullptr;
    std::vector<StencilDesc> host_sten;   // sizes for launch planning
    size_t bytes = 0;
    void release() {
        void *ps[] = {sten, entx, enty, ipsi, entic2, ents, e1, e2};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        sten = nullptr; entx = enty = ipsi = entic2 = nullptr; ents = nullptr; e1 = e2 = nullptr;
        mesh_id = 0; quirk = -1; geom = 0; n_sten = 0; host_sten.clear(); bytes = 0;
    }
};

// A fixed region of device memory handed out front to back (256-byte aligned): while an engine
// has one set, dbuf takes every buffer from it instead of allocating.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

struct pdds_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    std::mutex mu;
    int n_cu = 0;
    std::string name;
    std::vector<DevBuf> bufs;   // grow-only scratch, indexed by slot
    DevTables tables[2];        // two most recent (mesh, quirk) table sets
    int tables_next = 0;
    DevTables dom;              // one-stencil table of the rectangle point evaluations stop on
    double dom_rect[4] = {0, 0, 0, 0};
    int64_t record_budget = (int64_t)24 << 30;   // bytes of exit records per assembly batch
    int path_wgs_per_cu = 8;                      // persistent path-kernel work-groups per CU
    size_t lds_max = 64 * 1024;                   // LDS bytes a work-group may use
    Arena *arena = nullptr;     // set only inside pdds_point_eval_noalloc
    char *scratch = nullptr;    // device address of this device's no-allocation scratch array
};

struct pdds_lut {
    pdds_engine *e = nullptr;
    float *d = nullptr;
    LutView view{};
    std::vector<float> host;   // the device planes, for host evaluation (Dirichlet data of boundary knots)
    LutView hview{};           // view with host pointers into `host`
    int nx = 0, ny = 0, nf = 0, mode = 0;
};

uint64_t next_mesh_id();

struct pdds_mesh {
    Mesh m;
    uint64_t id;
    pdds_mesh(double a, double b, double c, double d, int n, int k, const double f[4])
        : m(a, b, c, d, n, k, f), id(next_mesh_id()) {}
    explicit pdds_mesh(Mesh &&mm) : m(std::move(mm)), id(next_mesh_id()) {}
};

enum BufSlot {
    B_KNOTS, B_STEN, B_ENTX, B_ENTY, B_ENTS, B_E1, B_E2, B_IPSI, B_ITEMS, B_STATS, B_PART, B_REC,
    B_KACC, B_COUNTER, B_PATCH, B_TABLES, B_G, B_GVAR, B_KSTAT, B_WREF, B_P0, B_P1, B_P2, B_P3, B_P4, B_P5, B_P6,
    B_P7, B_P8, B_P9, B_P10, B_P11, B_NSLOTS
};

pdds_status fail(pdds_engine *e, pdds_status s, const std::string &msg);

#define HIPCHK(expr)                                                                               \
    do {                                                                                           \
        hipError_t _r = (expr);                                                                    \
        if (_r != hipSuccess)                                                                      \
            return fail(e, PDDS_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_r));      \
    } while (0)

// grow-only device buffer for slot i
pdds_status dbuf(pdds_engine *e, int slot, size_t bytes, void **out);

template <class T>
pdds_status upload(pdds_engine *e, int slot, const std::vector<T> &v, T **out) {
    void *p;
    pdds_status s = dbuf(e, slot, v.size() * sizeof(T), &p);
    if (s) return s;
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, e->stream));
    *out = (T *)p;
    return PDDS_OK;
}

bool valid_pde(int pde);
pdds_status make_params(pdds_engine *e, const pdds_problem *p, const pdds_run *r, PdeParams &pp);
double host_g(const pdds_problem *p, double x, double y, const PdeParams &pp);
pdds_status check_run(pdds_engine *e, const pdds_run *r);

// One generic assembly over a list of knots sharing the stencil table.
struct AsmKnot {
    double x0, y0;
    uint32_t id;
    int stencil;
    int64_t n;
    int64_t out_off;
    int table = 0;          // index into the assembly's table list
    bool has_k2 = false;    // per-knot Helmholtz parameter (serving), else the problem's
    double k2 = 0.0;
    bool has_stream = false;// per-knot Philox stream (serving), else the run's
    uint32_t stream = 0;
    double h = 0.0;         // per-knot step (serving), else 0: the run's h or its adaptive_h rule
};

struct AsmOut {
    std::vector<KnotStat> kstats;
    double kernel_ms = 0, reduce_ms = 0;
    int64_t items = 0;
    int grid = 0, wg = 0, P = 0, S = 0;
};

int choose_P(int64_t n, int S);
pdds_status build_tables(pdds_engine *e, const std::vector<StencilDesc> &sten, const std::vector<double> &entx,
                         const std::vector<double> &enty, const std::vector<int8_t> &ents,
                         const std::vector<double> &entic2,
                         const std::vector<int32_t> &e1, const std::vector<int32_t> &e2,
                         const std::vector<double> &ipsi, DevTables &T);
// W_host (optional, n_vals values): per CSR value the reference interface value of its column's
// knot; the knots' equation means and variances then land in out.kstats[].Rm / Rv
pdds_status run_assembly(pdds_engine *e, const pdds_problem *prob, const pdds_run *r, const double glob[4],
                         const std::vector<const DevTables *> &tabs, const std::vector<AsmKnot> &knots,
                         int64_t n_vals, double *G_host, double *Gvar_host, AsmOut &out,
                         const double *W_host = nullptr);
void mesh_tables(const Mesh &m, int quirk, std::vector<StencilDesc> &sten, std::vector<double> &entx,
                 std::vector<double> &enty, std::vector<int8_t> &ents, std::vector<double> &entic2,
                 std::vector<int32_t> &e1,
                 std::vector<int32_t> &e2, std::vector<double> &ipsi);
pdds_status mesh_dev_tables(pdds_engine *e, const pdds_mesh *m, int quirk, DevTables **out);
void fill_counters(pdds_counters *ctr, const AsmOut &o, double total_ms);
pdds_status solve_system(pdds_engine *e, int32_t n, const int64_t *row_ptr, const int32_t *cols, const double *G,
                         const double *B, double *u, double tol, int32_t maxit, int32_t *iters, double *relres);
double now_ms();
