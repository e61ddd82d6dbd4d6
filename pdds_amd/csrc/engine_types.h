// engine_types.h -- device-resident descriptors shared by host and kernels.
#pragma once
#include <stdint.h>
#include "pdds_pde.h"

namespace pdds {

constexpr int kNBG = 8;                    // sub-batches reduced per pass over Psi^-1
constexpr int kReduceWG = 256;            // reduction work-group
constexpr int kRbfChunk = 16;             // side entries per direct exponential (rectangle sides)

// LDS of the reduction kernel: Phi [ne_cap][kNBG], accS/accQ [nnz_cap] (FP64), the group's
// records x, y, Y and score [kNBG*S] (FP64) and per-(sub-batch, side) path lists [kNBG][4][S]
// (uint16)
inline size_t reduce_lds_bytes(int ne_cap, int nnz_cap, int S) {
    const size_t g = (size_t)kNBG * S;
    size_t b = ((size_t)kNBG * ne_cap + 2 * (size_t)nnz_cap + 4 * g) * sizeof(double);
    b += g * 4 * sizeof(uint16_t);
    return (b + 15) & ~(size_t)15;
}
// FP32 exit records carry the exit side + 1 (0 = none, 1..5 = sides -1..4) in the low bits of
// Y: a float widened to double leaves the low 29 bits of its mantissa zero.
constexpr uint64_t kSideBitsMask = 7ULL;
constexpr int kMaxQueues = 8;              // path queues: one per XCD of an MI355X (8 XCDs x 32 CUs)

// One Monte Carlo knot of an assembly batch.
struct KnotDesc {
    double x0, y0;          // start point
    uint32_t knot_id;       // Philox counter word 2 (knot index within its request / mesh)
    uint32_t stream;        // Philox key word 1 (the run's stream, or the serving request's)
    int32_t table;          // stencil table of the knot (TableDesc index of the launch)
    int32_t stencil;        // index into that table's stencils
    int64_t n_paths;
    int64_t path_off;       // global index of the knot's first path in the batch (ascending)
    int64_t out_off;        // first value of the knot's G row in the CSR value array
    int64_t item_begin;     // first work item of the knot
    int64_t item_count;
    double k2;              // Helmholtz parameter of the knot's problem
    double h, sqrth;        // Euler-Maruyama step of the knot's paths (the run's, or its adaptive h)
};

// One distinct stencil: stopping rectangle, four side lists (N,S,E,W), Psi^-1 per side,
// and for each CSR column of the row up to two contributing side entries.
struct StencilDesc {
    double rect[4];         // rectangular stopping domain x0, y0, x1, y1
    double geo[6];          // annular-sector stopping domain r0, r1, cos/sin ray a, cos/sin ray b
    double bbox[4];         // bounding box of the stopping domain (host: LUT tiles)
    double inv_c2[4];       // 1 / c^2 of the Gaussian shape c per side
    int64_t ipsi_off[4];    // row-major m x m blocks in the ipsi array
    int32_t msz[4];
    int32_t ent_off;        // first side entry in ent_x / ent_y / ent_side
    int32_t csr_nnz;        // columns in the CSR row
    int32_t csr_off;        // first column map entry in csr_e1 / csr_e2
    int32_t gmask;          // bit s set: side s lies on the global boundary
};

// The device stencil tables of one mesh: a launch may carry knots of several meshes (serving),
// each knot naming its table.
struct TableDesc {
    const StencilDesc *sten;
    const double *entx, *enty;     // side-entry positions
    const int8_t *ents;            // side of each entry
    const double *entic2;          // 1 / c^2 of each entry's Gaussian (its side's c, or its own)
    const int32_t *e1, *e2;        // contributing entries of each CSR column
    const double *ipsi;            // Psi^-1 blocks
};

struct ItemDesc {
    int32_t knot;           // index into the knot table
    int32_t n;              // paths in the item
    int64_t p_begin;        // first path index within the knot
    int64_t part_off;       // 2*nnz partials (sum, sum of squares / n_beta)
};

struct ItemStat {
    double Bs, BBs;
    double Rs, RQs;         // equation residual sums (ReduceArgs::wref set): sum R_b, sum R_b^2 / n_b
    uint32_t cnt[6];        // exits per side 0..4, and paths stopped inside (no side)
};

struct KnotStat {
    double B, BB;
    double Rm, Rv;          // mean and per-path variance of the knot's equation estimator
    unsigned long long steps;
    unsigned long long capped;
    unsigned long long cnt[6];
};

// Per-path exit records, indexed by global path index: the FP64 exit values as one aligned
// 32-byte record (a lane stores it with two 16-byte writes when its path stops, so a record
// reaches memory as one 32-byte piece).  The exit side is not stored for FP64 paths: the
// reduction classifies the exit point again, with the same test on the same doubles.  Step
// counts are not stored per path: a lane sums the steps and step-cap stops of its stopped paths
// in registers and adds them to its knot's counters kacc[2k] (steps) and kacc[2k + 1] (capped)
// when it moves to a path of another knot and when it ends.
struct alignas(32) ExitRec {
    double x, y, Y, score;
};
struct Records {
    ExitRec *r;
    unsigned long long *kacc;   // 2 per knot of the launch, zero before the path kernel
};

// A patch: knots whose LUT tile (the plane-a texels their stencil rectangle can touch) is the
// same; their paths are [pbeg, pend) of the batch.  tiled == 0: the tile does not fit in LDS
// and the patch reads the LUT from global memory.
struct PatchDesc {
    int64_t pbeg, pend;
    int32_t c0, r0, tw, th;
    int32_t tiled, pitch;   // pitch: texels per LDS tile row (tw plus the run's lut_pad)
};

struct PathKernelArgs {
    const KnotDesc *knots;
    int32_t n_knots;
    const TableDesc *tables;
    int64_t total;          // paths in the batch
    int64_t base;           // plain launches: first path of this launch
    // persistent launches: queue q serves paths [q*total/n_queues, (q+1)*total/n_queues)
    // through counter[q]; work-group b starts on queue b % n_queues (the XCD it is
    // dispatched to) and moves to the next queue when its own is drained
    unsigned long long *counter;
    int32_t n_queues;
    int64_t pool_chunk;     // paths a wave takes from a queue at once (64 x paths per lane)
    // LDS-tiled launches: one queue per patch (counter[p] counts paths handed out of patch p)
    const struct PatchDesc *patches;
    int32_t n_patches;
    Records rec;
    double glob[4];
    double h, sqrth, beta;
    uint32_t seed, stream;
    int64_t max_steps;
    PdeParams pp;
};

struct ReduceArgs {
    const KnotDesc *knots;
    const TableDesc *tables;
    const ItemDesc *items;
    int64_t item_base;
    Records rec;
    ItemStat *stats;
    double *partials;
    // optional: per CSR value (out_off + c), the reference interface value u_ref of the column's
    // knot; the reduction then also sums R_b = B_b - sum_c G_b,c u_ref per sub-batch b
    const double *wref;
    int32_t ne_cap, nnz_cap;
    int32_t S;
    // 1: every stencil side is straight and uniformly spaced with one Gaussian width (rectangle
    // meshes), and the RBF values of a side come from rbf_side_chunk
    int32_t side_rbf;
    // exit sides (ReduceSide): from the low bits of the records' Y (FP32 paths), or by
    // classifying the exit point as write_exit does for FP64 paths, on the global
    // rectangle of centre and half-sides gc (classify_coord_c) or the disc of radius gc[0]
    int32_t side_mode;
    double gc[4];
};
enum ReduceSide { SIDE_BITS = 0, SIDE_RECT = 1, SIDE_DISC = 2 };

struct FinalizeArgs {
    const KnotDesc *knots;
    const TableDesc *tables;
    const ItemDesc *items;
    const ItemStat *stats;
    const double *partials;
    int32_t n_knots;
    int32_t S;
    double *G, *Gvar;
    KnotStat *kstats;
    const unsigned long long *kacc;   // steps and capped paths per knot (Records::kacc)
};

struct PathsArgs {
    double x0, y0;
    double rect[4];
    uint32_t knot, seed, stream;
    int32_t phase2;
    int64_t path0, n;
    double h, sqrth, beta;
    int64_t max_steps;
    PdeParams pp;
    double *xt, *yt, *Y, *Z, *xi;
    int64_t *steps;
    double *xt2, *yt2, *Y2, *Z2, *xi2;
    int64_t *steps2;
};

// Launchers (defined in engine_f64.hip / engine_f32.hip)
hipError_t launch_path_tiled_f64(int geom, bool mixed, int grid, size_t lds, hipStream_t s, const PathKernelArgs &a);
hipError_t launch_path_tiled_f32(int geom, bool mixed, int grid, size_t lds, hipStream_t s, const PathKernelArgs &a);
int path_tiled_wg(int precision, int step, int geom);   // block size of the LDS-tiled kernel
hipError_t launch_path_f64(int pde, int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                           const PathKernelArgs &a);
hipError_t launch_path_f32(int pde, int geom, bool persistent, bool mixed, int grid, int wg, hipStream_t s,
                           const PathKernelArgs &a);
hipError_t launch_reduce(int grid, int wg, size_t shmem, hipStream_t s, const ReduceArgs &a);
hipError_t launch_paths_f64(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a);
hipError_t launch_paths_f32(int pde, int grid, int wg, hipStream_t s, const PathsArgs &a);
hipError_t launch_finalize(int grid, int wg, hipStream_t s, const FinalizeArgs &f);
hipError_t launch_lut_eval(const LutView &L, int field, int64_t n, const double *x, const double *y, double *out,
                           hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, int64_t n, unsigned long long v, hipStream_t s);

}  // namespace pdds
