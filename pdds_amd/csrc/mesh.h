// mesh.h -- host-side interface/knot/stencil bookkeeping of PDDSparse.
//
// C++ restatement of PDDSparseGM's geometry for the main_EXTENDED
// configuration (INTERSECTIONS_YES, STEN_ELONG = 0.5, square grids):
//   Fullfill_interfaces     PDDAlgorithms/PDDSparseGM.cpp:346-689
//   Solve's knot dispatch   PDDAlgorithms/PDDSparseGM.cpp:705-756
//   Node_Position           PDDAlgorithms/PDDSparseGM.cpp:3156-3167
//   Interface::N/S/E/W      Meshes/interface.cpp:16-93
//   Labels_Stencil          PDDAlgorithms/PDDSparseGM.cpp:1330-1433
//   Compute_Stencil         PDDAlgorithms/PDDSparseGM.cpp:1845-2041
//   Stencil::Init           Meshes/stencil.cpp:11-85
//   Stencil::Compute_ipsi   Meshes/stencil.cpp:115-243 (full-pivot LU inverse, Frobenius cond)
// Knots sharing an identical stencil (all interior knots of a segment) share one
// Psi^-1, computed once.
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "pdds_math.h"

namespace pdds {

enum Dir { North = 0, South = 1, East = 2, West = 3 };

struct Interface {
    std::array<int, 2> label;
    std::vector<int> index;
    std::vector<double> px, py;
    bool In(int k) const;
    std::vector<std::array<int, 2>> N() const;
    std::vector<std::array<int, 2>> S() const;
    std::vector<std::array<int, 2>> E() const;
    std::vector<std::array<int, 2>> W() const;
};

struct Stencil {
    double rect[4];                        // rectangle x0, y0, x1, y1 (rectangular meshes)
    // annular sector r0, r1, cos/sin of the start ray a, cos/sin of the end ray b (disc meshes)
    double geo[6] = {0, 0, 0, 0, 0, 0};
    double bbox[4] = {0, 0, 0, 0};         // bounding box of the stopping domain (LUT tiles)
    int kind[4];
    std::vector<int> idx[4];
    std::vector<double> px[4], py[4];
    // derived
    double cshape[4] = {1, 1, 1, 1};
    std::vector<double> cent[4];           // per-entry Gaussian widths (empty: cshape of the side)
    double cond[4] = {0, 0, 0, 0};
    std::vector<double> ipsi[4];          // row-major m x m
    // distinct-side index within the mesh: sides with the same index have the same entries,
    // widths and Psi^-1 bit for bit (-1: empty side)
    int side_id[4] = {-1, -1, -1, -1};
    std::vector<int> cols;                 // sorted unique CSR columns
    std::vector<int> e1, e2;               // per column: contributing side entries (concatenated N,S,E,W), -1 none
    std::vector<int> e1_sum, e2_sum;       // same, summing shared corners
    int gmask = 0;
    bool singular = false;
    // > 0: shrink a side's RBF shape by 0.8 until its Psi is invertible with cond below this
    double cond_max = 0.0;
    int total() const { return (int)(idx[0].size() + idx[1].size() + idx[2].size() + idx[3].size()); }
    bool same_as(const Stencil &o) const;
};

class Mesh {
  public:
    Mesh(double sw_x, double sw_y, double ne_x, double ne_y, int n_sub, int n_node, const double fac[4]);
    // Disc of the given radius cut into n_rings rings x n_sectors sectors; each interface (arc or
    // radial segment) carries a number of knots drawn uniformly from [kmin, kmax] with `seed`.
    static Mesh disc(double radius, int n_rings, int n_sectors, uint32_t seed, int kmin, int kmax, double shape);
    int geom = GeomRect;
    double radius = 0.0;                    // disc meshes
    int n_knots() const { return nknots_; }
    const std::vector<Interface> &interfaces() const { return ifs_; }
    double glob[4];
    std::vector<double> x0, y0, npx, npy;   // start / reported positions
    std::vector<int> on_boundary;
    std::vector<int> stencil_of;            // knot -> distinct stencil id
    std::vector<Stencil> stencils;
    std::vector<int64_t> row_ptr;           // CSR of G (no identity); boundary knots: empty rows
    std::vector<int> cols;
    bool any_singular = false;
    double fac[4];

  private:
    Mesh() = default;
    // dedup per-knot stencils (cshape and gmask set), Psi^-1 per side, CSR rows, column maps
    void finalize(std::vector<Stencil> &per_knot, const std::vector<char> &has_stencil);
    int find(int a, int b) const;
    std::map<Dir, std::vector<std::array<int, 2>>> labels_stencil(int k) const;
    Stencil compute_stencil(int k) const;
    int n_ = 0, k_ = 0;
    int nknots_ = 0;
    std::vector<Interface> ifs_;
    std::map<std::array<int, 2>, int> imap_;
};

// Full-pivot LU inverse of the Gaussian RBF matrix; returns false if singular.
// With cw, centre j has its own width cw[j]: Psi_ij = exp(-|p_i - p_j|^2 / cw[j]^2).
bool psi_inverse(const std::vector<double> &px, const std::vector<double> &py, double c, std::vector<double> &ipsi,
                 double &cond, const std::vector<double> *cw = nullptr);

}  // namespace pdds
