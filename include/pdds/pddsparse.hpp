// pddsparse.hpp -- C++ host API over the C ABI of libpdds.so, with the names of the
// reference's plain PDDSparse path so that a driver written against it changes little:
//   struct bvp                      BVPs/BVP.hpp:89-137 (here: a compiled-in PDE id, optional LUT,
//                                   and a host function for the analytic solution written to
//                                   solution.csv)
//   class EMFKACSolver              Solvers/EMFKACSolver.hpp:48-764
//     Solve_OMP_Analytic(...)       :711-729 (stencil overload) and :686-691 (point overload)
//   class PDDSparseGM               PDDAlgorithms/PDDSparseGM.{hpp,cpp}
//     PDDSparseGM(argc, argv, file) :85-97  (ReadFromFile :99-345, Fullfill_interfaces :346-689)
//     Solve(bvp)                    :691-851 (all knots in one device batch)
//     Compute_Solution(bvp)         :3168-3246 (device BiCGSTAB; G/B/variance/solution CSVs)
// Errors throw pdds::Error (the reference prints and continues); nothing calls exit().
#pragma once
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../pdds_engine.h"

namespace pdds {

struct Error : std::runtime_error {
    pdds_status status;
    Error(pdds_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(pdds_status s, const pdds_engine *e, const char *what) {
    if (s != PDDS_OK)
        throw Error(s, std::string(what) + ": " + pdds_status_string(s) + (e ? std::string(" (") + pdds_last_error(e) + ")" : ""));
}

// Problem descriptor: which coefficient functor runs on the device.
struct bvp {
    int32_t pde = PDDS_PDE_MONEGROS;
    double k2 = 0.0;
    const pdds_lut *lut = nullptr;
    int32_t neumann = 0;     // reflecting global sides (PDDS_SIDE_* bits), the reference's bvp.Neumann
    double varphi = 0.0;     // Robin coefficient on those sides (<= 0)
    double rho = 0.0;        // Lepingle step within this distance of a reflecting side (0: always)
    const pdds_lut *ulut = nullptr;   // PDDS_PDE_TABLE: u, the Dirichlet data (1 field)
    const pdds_lut *vlut = nullptr;   // PDDS_STEP_VRCV: mu1, mu2, F1, F2
    std::function<double(double, double)> u;   // analytic solution (solution.csv); optional
    pdds_problem desc() const { pdds_problem p{pde, k2, lut, neumann, varphi, rho, ulut, vlut}; return p; }
};

// Owns one engine (one device).
class Device {
  public:
    explicit Device(int device = 0) { check(pdds_engine_create(device, &e_), nullptr, "pdds_engine_create"); }
    ~Device() { pdds_engine_destroy(e_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    pdds_engine *get() const { return e_; }

  private:
    pdds_engine *e_ = nullptr;
};

// Per-knot solver with the reference's entry points.
class EMFKACSolver {
  public:
    explicit EMFKACSolver(Device &d, uint32_t seed = 42) : dev_(d) {
        pdds_run_defaults(&run);
        run.seed = seed;
    }
    pdds_run run;
    pdds_counters last{};

    // Stencil overload: G row (columns G_j ascending), per-path variances, B and E[score^2].
    // The knot's stencil comes from the mesh (Compute_Stencil); rho and c2 of the reference's
    // signature are kept for source compatibility (rho is unused for Dirichlet problems, the
    // shapes come from the mesh's fac).
    void Solve_OMP_Analytic(const pdds_mesh *mesh, int knot, unsigned N, double h, double /*rho*/, const bvp &p,
                            std::vector<double> &G, std::vector<double> &G_var, std::vector<int> &G_j, double &B,
                            double &BB) {
        run.h = h;
        int64_t nnz = 0;
        std::vector<int64_t> rp(pdds_mesh_n_knots(mesh) + 1);
        check(pdds_mesh_csr(mesh, rp.data(), nullptr, &nnz), dev_.get(), "pdds_mesh_csr");
        std::vector<int32_t> cols(nnz > 0 ? nnz : 1);
        check(pdds_mesh_csr(mesh, rp.data(), cols.data(), &nnz), dev_.get(), "pdds_mesh_csr");
        const int64_t n = rp[knot + 1] - rp[knot];
        G.assign(n, 0.0);
        G_var.assign(n, 0.0);
        G_j.assign(cols.begin() + rp[knot], cols.begin() + rp[knot + 1]);
        pdds_problem d = p.desc();
        check(pdds_solve_knot(dev_.get(), mesh, &d, &run, knot, N, G.data(), G_var.data(), &B, &BB, &last),
              dev_.get(), "pdds_solve_knot");
    }

    // Point overload: u(X0) on the rectangle boundary_parameters = {x0, y0, x1, y1}.
    double Solve_OMP_Analytic(double x0, double y0, unsigned N, double h, double /*rho*/, const bvp &p,
                              const double boundary_parameters[4], double *var = nullptr) {
        run.h = h;
        pdds_problem d = p.desc();
        int64_t n = N;
        double u = 0.0, v = 0.0;
        check(pdds_point_eval(dev_.get(), &d, &run, boundary_parameters, 1, &x0, &y0, &n, &u, &v, &last),
              dev_.get(), "pdds_point_eval");
        if (var) *var = v;
        return u;
    }

  private:
    Device &dev_;
};

// The plain PDDSparse driver.
class PDDSparseGM {
  public:
    double h0 = 0.005, eps = 0.5, T_start = INFINITY;
    double fac[4] = {1.0, 1.0, 1.0, 1.0};
    double SW[2] = {-1.0, -1.0}, NE[2] = {1.0, 1.0};
    unsigned N = 1000;
    int iN = 2, nN = 8;
    uint32_t seed = 42;
    std::string out_dir = "Output";
    pdds_run run;
    pdds_counters counters{};
    std::vector<double> G, G_var, B, BB, u;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> cols;
    int bicgstab_iters = 0;
    double bicgstab_relres = 0.0;

    PDDSparseGM(int /*argc*/, char ** /*argv*/, const std::string &file, Device &dev) : dev_(dev) {
        t0_ = now();
        ReadFromFile(file);
        Update_TimeFile("Reading configuration File", 1);
        pdds_run_defaults(&run);
        run.seed = seed;
        check(pdds_mesh_create(SW[0], SW[1], NE[0], NE[1], iN, nN, fac, &mesh_), dev_.get(), "pdds_mesh_create");
        Update_TimeFile("Fullfilling interfaces", 1);
    }
    ~PDDSparseGM() { pdds_mesh_destroy(mesh_); }
    PDDSparseGM(const PDDSparseGM &) = delete;
    PDDSparseGM &operator=(const PDDSparseGM &) = delete;

    const pdds_mesh *mesh() const { return mesh_; }
    int nNodes() const { return pdds_mesh_n_knots(mesh_); }

    // configuration.txt: whitespace-separated "key= value" tokens (PDDAlgorithms/PDDSparseGM.cpp:99-345);
    // "key = value" and "key=value" are accepted too.  N_subdo / N_node set both directions.
    void ReadFromFile(const std::string &file) {
        std::ifstream in(file);
        if (!in) throw Error(PDDS_ERR_INVALID, "cannot read configuration file " + file);
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '%') continue;
            std::string flat;
            for (char c : line) flat += (c == '=') ? std::string(" = ") : std::string(1, c);
            std::istringstream ss(flat);
            std::string key, eq;
            if (!(ss >> key >> eq) || eq != "=") continue;
            std::vector<std::string> v;
            std::string t;
            while (ss >> t) v.push_back(t);
            auto num = [&](size_t i) { return i < v.size() ? std::stod(v[i]) : 0.0; };
            if (key == "h0") h0 = num(0);
            else if (key == "eps") eps = num(0);
            else if (key == "T") T_start = num(0);
            else if (key == "fac") for (int i = 0; i < 4; ++i) fac[i] = num(i);
            else if (key == "SW") { SW[0] = num(0); SW[1] = num(1); }
            else if (key == "NE") { NE[0] = num(0); NE[1] = num(1); }
            else if (key == "N") N = (unsigned)num(0);
            else if (key == "N_subdo") iN = (int)num(0);
            else if (key == "N_node") nN = (int)num(0);
            else if (key == "seed") seed = (uint32_t)num(0);
        }
    }

    // Monte Carlo assembly of G and B over every knot (the worker loop, one device batch).
    void Solve(const bvp &p) {
        run.h = h0;
        int64_t nnz = 0;
        row_ptr.assign(nNodes() + 1, 0);
        check(pdds_mesh_csr(mesh_, row_ptr.data(), nullptr, &nnz), dev_.get(), "pdds_mesh_csr");
        cols.assign(nnz > 0 ? nnz : 1, 0);
        check(pdds_mesh_csr(mesh_, row_ptr.data(), cols.data(), &nnz), dev_.get(), "pdds_mesh_csr");
        cols.resize(nnz);
        G.assign(nnz, 0.0);
        G_var.assign(nnz, 0.0);
        B.assign(nNodes(), 0.0);
        BB.assign(nNodes(), 0.0);
        pdds_problem d = p.desc();
        check(pdds_assemble(dev_.get(), mesh_, &d, &run, nullptr, N, G.data(), G_var.data(), B.data(), BB.data(),
                            &counters),
              dev_.get(), "pdds_assemble");
        Update_TimeFile("MC Simulations", 1);
        Compute_Solution(p);
    }

    // (I + G) u = B on the device; writes the reference's CSVs.
    void Compute_Solution(const bvp &p) {
        const int n = nNodes();
        u.assign(n, 0.0);
        write_debug_csvs();
        Update_TimeFile("Building Up G and B", 1);
        check(pdds_solve(dev_.get(), n, row_ptr.data(), cols.data(), G.data(), B.data(), u.data(), 1e-14, 0,
                         &bicgstab_iters, &bicgstab_relres),
              dev_.get(), "pdds_solve");
        Update_TimeFile("Solving system", 1);
        std::vector<double> x0(n), y0(n), px(n), py(n);
        check(pdds_mesh_knots(mesh_, x0.data(), y0.data(), px.data(), py.data(), nullptr, nullptr), dev_.get(),
              "pdds_mesh_knots");
        FILE *f = std::fopen((out_dir + "/solution.csv").c_str(), "w");
        if (!f) throw Error(PDDS_ERR_INVALID, "cannot write " + out_dir + "/solution.csv");
        std::fprintf(f, "Knot_index,x,y,sol_analytic,sol_PDDS,err,rerr\n");
        for (int i = 0; i < n; ++i) {
            double sol = p.u ? p.u(px[i], py[i]) : NAN;
            double err = u[i] - sol;
            std::fprintf(f, "%d,%e,%e,%e,%e,%e,%e\n", i, px[i], py[i], sol, u[i], err, std::fabs(err) / sol);
        }
        std::fclose(f);
        Update_TimeFile("Writing Solution", 1);
    }

    // task,total,task_time,N_process lines (PDDAlgorithms/PDDSparseGM.cpp:4397-4404)
    void Update_TimeFile(const std::string &task, int nproc) {
        double t = now();
        FILE *f = std::fopen((out_dir + "/Debug/times.txt").c_str(), "a");
        if (f) {
            std::fprintf(f, "%s,%e,%e,%d\n", task.c_str(), t - t0_, t - tprev_, nproc);
            std::fclose(f);
        }
        tprev_ = t;
    }

  private:
    Device &dev_;
    pdds_mesh *mesh_ = nullptr;
    double t0_ = 0.0, tprev_ = 0.0;

    double now() {
        double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        if (tprev_ == 0.0) tprev_ = t;
        return t;
    }

    // Column-major triplet dumps as Eigen writes them (outer index = column, rows ascending).
    void write_debug_csvs() {
        const int n = nNodes();
        std::vector<std::vector<std::pair<int, double>>> byc(n), byc_var(n);
        for (int r = 0; r < n; ++r)
            for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                byc[cols[k]].push_back({r, G[k]});
                byc_var[cols[k]].push_back({r, G_var[k]});
            }
        auto dump = [&](const std::string &name, const char *hdr, bool add_identity,
                        const std::vector<std::vector<std::pair<int, double>>> &m) {
            FILE *f = std::fopen((out_dir + "/Debug/" + name).c_str(), "w");
            if (!f) throw Error(PDDS_ERR_INVALID, "cannot write " + out_dir + "/Debug/" + name);
            std::fprintf(f, "%s\n", hdr);
            for (int c = 0; c < n; ++c) {
                std::map<int, double> col;
                for (auto &e : m[c]) col[e.first] += e.second;
                if (add_identity) col[c] += 1.0;
                for (auto &e : col) std::fprintf(f, "%d,%d,%e\n", e.first, c, e.second);
            }
            std::fclose(f);
        };
        dump("G_var.csv", "Gvar_i,Gvar_j,Gvar_ij", false, byc_var);
        dump("G.csv", "G_i,G_j,G_ij", true, byc);
        FILE *f = std::fopen((out_dir + "/Debug/B_var.csv").c_str(), "w");
        if (!f) throw Error(PDDS_ERR_INVALID, "cannot write B_var.csv");
        std::fprintf(f, "B_i,B_j,B_ij\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%d,%d,%e\n", i, 0, B[i] * B[i] - BB[i]);  // B^2 - E[s^2]
        std::fclose(f);
        f = std::fopen((out_dir + "/Debug/B.csv").c_str(), "w");
        if (!f) throw Error(PDDS_ERR_INVALID, "cannot write B.csv");
        std::fprintf(f, "B_i,B_j,B_ij\n");
        for (int i = 0; i < n; ++i) std::fprintf(f, "%d,%d,%e\n", i, 0, B[i]);
        std::fclose(f);
    }
};

}  // namespace pdds
