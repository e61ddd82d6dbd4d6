// mesh.cpp -- host-side interface/knot/stencil bookkeeping (see mesh.h).
#include "mesh.h"

#include <algorithm>
#include <atomic>
#include <map>
#include <thread>
#include <cmath>

#include "pdds_math.h"

namespace pdds {

bool Interface::In(int k) const {
    for (int v : index)
        if (v == k) return true;
    return false;
}
// Meshes/interface.cpp:16-93 (vertical interfaces have even label[1])
std::vector<std::array<int, 2>> Interface::N() const {
    if (label[1] % 2 == 0) return {{label[0], label[1] + 1}, {label[0] + 1, label[1] + 1}};
    return {{label[0], label[1] + 2}};
}
std::vector<std::array<int, 2>> Interface::S() const {
    if (label[1] % 2 == 0) return {{label[0], label[1] - 1}, {label[0] + 1, label[1] - 1}};
    return {{label[0], label[1] - 2}};
}
std::vector<std::array<int, 2>> Interface::E() const {
    if (label[1] % 2 == 0) return {{label[0] + 1, label[1]}};
    return {{label[0], label[1] - 1}, {label[0], label[1] + 1}};
}
std::vector<std::array<int, 2>> Interface::W() const {
    if (label[1] % 2 == 0) return {{label[0] - 1, label[1]}};
    return {{label[0] - 1, label[1] - 1}, {label[0] - 1, label[1] + 1}};
}

bool Stencil::same_as(const Stencil &o) const {
    for (int i = 0; i < 4; ++i) {
        if (rect[i] != o.rect[i] || kind[i] != o.kind[i] || idx[i] != o.idx[i] || px[i] != o.px[i] ||
            py[i] != o.py[i])
            return false;
    }
    return true;
}

int Mesh::find(int a, int b) const {
    auto it = imap_.find({a, b});
    return it == imap_.end() ? -1 : it->second;
}

Mesh::Mesh(double sw_x, double sw_y, double ne_x, double ne_y, int n_sub, int n_node, const double facv[4])
    : n_(n_sub), k_(n_node) {
    glob[0] = sw_x; glob[1] = sw_y; glob[2] = ne_x; glob[3] = ne_y;
    for (int i = 0; i < 4; ++i) fac[i] = facv[i];
    const int n = n_, k = k_;
    const int nif = 2 * n * (n - 1);
    ifs_.resize(nif);
    std::vector<double> stx(nif), sty(nif), enx(nif), eny(nif);
    // Fullfill_interfaces: segment end points from accumulated long increments
    const double L0x = (ne_x - sw_x) / (1.0 * n), L1y = (ne_y - sw_y) / (1.0 * n);
    double P0x = sw_x + 0.0, P0y = sw_y + L1y, P1x = sw_x + L0x, P1y = sw_y + 0.0;
    int cen = 0;
    for (int i = 0; i < n - 1; ++i) {
        for (int j = 0; j < n; ++j) {
            stx[cen] = P0x + L0x * j; sty[cen] = P0y + 0.0 * j;
            enx[cen] = P0x + L0x * (j + 1); eny[cen] = P0y + 0.0 * (j + 1);
            ifs_[cen].label = {j, 2 * i + 1};
            ++cen;
            stx[cen] = P1x + 0.0 * j; sty[cen] = P1y + L1y * j;
            enx[cen] = P1x + 0.0 * (j + 1); eny[cen] = P1y + L1y * (j + 1);
            ifs_[cen].label = {i, 2 * j};
            ++cen;
        }
        P0x = P0x + 0.0; P0y = P0y + L1y;
        P1x = P1x + L0x; P1y = P1y + 0.0;
    }
    for (int i = 0; i < nif; ++i) imap_[ifs_[i].label] = i;
    // knot numbering: vertical lines first, crossings shared with the horizontal rows
    std::vector<int> cross((n - 1) * (n - 1) + 1);
    int cent = 0, cc = 0;
    for (int i = 0; i < n - 1; ++i) {
        for (int j = 0; j <= 2 * n - 2; j += 2) {
            Interface &f = ifs_[find(i, j)];
            f.index.push_back(cent);
            if (j != 0) cross[(j / 2 - 1) * (n - 1) + i] = cent;
            for (int q = 0; q < k - 1; ++q) f.index.push_back(++cent);
            if (j == 2 * n - 2) ++cent;
        }
    }
    for (int j = 1; j <= 2 * n - 3; j += 2) {
        for (int i = 0; i < n; ++i) {
            Interface &f = ifs_[find(i, j)];
            if (i != 0) f.index.push_back(cross[cc++]);
            else f.index.push_back(cent++);
            for (int q = 1; q < k - 1; ++q) f.index.push_back(cent++);
            if (i != n - 1) f.index.push_back(cross[cc]);
            else f.index.push_back(cent++);
        }
    }
    nknots_ = cent;
    for (int i = 0; i < nif; ++i) {
        Interface &f = ifs_[i];
        const int m = (int)f.index.size();
        const double ix = (enx[i] - stx[i]) / (double)(m - 1), iy = (eny[i] - sty[i]) / (double)(m - 1);
        double vx = stx[i], vy = sty[i];
        f.px.resize(m);
        f.py.resize(m);
        for (int q = 0; q < m; ++q) { f.px[q] = vx; f.py[q] = vy; vx = vx + ix; vy = vy + iy; }
    }
    // start points in the server's dispatch order; reported positions = last interface
    x0.assign(nknots_, 0.0); y0.assign(nknots_, 0.0);
    npx.assign(nknots_, 0.0); npy.assign(nknots_, 0.0);
    for (int Iy = 0; Iy <= 2 * n - 2; ++Iy) {
        if (Iy % 2 == 0) {
            for (int Ix = 0; Ix <= n - 2; ++Ix) {
                const Interface &f = ifs_[find(Ix, Iy)];
                for (size_t q = (Iy == 0 ? 0 : 1); q < f.index.size(); ++q) { x0[f.index[q]] = f.px[q]; y0[f.index[q]] = f.py[q]; }
            }
        } else {
            for (int Ix = 0; Ix <= n - 1; ++Ix) {
                const Interface &f = ifs_[find(Ix, Iy)];
                size_t lo = (Ix == 0) ? 0 : 1, hi = (Ix == n - 1) ? f.index.size() : f.index.size() - 1;
                for (size_t q = lo; q < hi; ++q) { x0[f.index[q]] = f.px[q]; y0[f.index[q]] = f.py[q]; }
            }
        }
    }
    for (const Interface &f : ifs_)
        for (size_t q = 0; q < f.index.size(); ++q) { npx[f.index[q]] = f.px[q]; npy[f.index[q]] = f.py[q]; }
    on_boundary.assign(nknots_, 0);
    for (int q = 0; q < nknots_; ++q) {
        double ex, ey, nx, ny;
        int kd;
        on_boundary[q] = std::fabs(rect_dist<double>(glob, x0[q], y0[q], ex, ey, nx, ny, kd)) <= 1e-8;
    }
    // stencils of every knot (independent, spread over host threads)
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nth_all = std::min<size_t>((size_t)(hw ? hw : 1), 16);
    std::vector<Stencil> per_knot(nknots_);
    {
        std::atomic<int> nextq{0};
        auto work = [&]() {
            for (int q; (q = nextq.fetch_add(1)) < nknots_;) per_knot[q] = compute_stencil(q);
        };
        std::vector<std::thread> pool;
        for (size_t i = 1; i < std::min<size_t>(nth_all, (size_t)nknots_ / 64 + 1); ++i) pool.emplace_back(work);
        work();
        for (std::thread &th : pool) th.join();
    }
    for (Stencil &s : per_knot) {
        for (int d = 0; d < 4; ++d) s.cshape[d] = fac[s.kind[d] < 0 ? 0 : (s.kind[d] > 3 ? 3 : s.kind[d])];
        for (int i = 0; i < 4; ++i) s.bbox[i] = s.rect[i];
        if (std::fabs(s.rect[0] - glob[0]) < 1e-8) s.gmask |= 1 << West;
        if (std::fabs(s.rect[1] - glob[1]) < 1e-8) s.gmask |= 1 << South;
        if (std::fabs(s.rect[2] - glob[2]) < 1e-8) s.gmask |= 1 << East;
        if (std::fabs(s.rect[3] - glob[3]) < 1e-8) s.gmask |= 1 << North;
    }
    std::vector<char> has(nknots_, 1);
    finalize(per_knot, has);
}

void Mesh::finalize(std::vector<Stencil> &per_knot, const std::vector<char> &has_stencil) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nth_all = std::min<size_t>((size_t)(hw ? hw : 1), 16);
    // deduplicated in knot order: a knot takes the first earlier stencil equal to its own
    stencil_of.assign(nknots_, -1);
    std::map<std::vector<int64_t>, int> seen;
    for (int q = 0; q < nknots_; ++q) {
        if (!has_stencil[q]) continue;
        const Stencil &s = per_knot[q];
        std::vector<int64_t> key;
        for (int i = 0; i < 4; ++i) {
            key.push_back((int64_t)f64_bits(s.rect[i]));
            key.push_back((int64_t)f64_bits(s.cshape[i]));
            key.push_back(s.kind[i]);
            key.push_back((int64_t)s.idx[i].size());
            key.insert(key.end(), s.idx[i].begin(), s.idx[i].end());
        }
        for (int i = 0; i < 6; ++i) key.push_back((int64_t)f64_bits(s.geo[i]));
        auto it = seen.find(key);
        if (it == seen.end()) {
            it = seen.emplace(std::move(key), (int)stencils.size()).first;
            stencils.push_back(std::move(per_knot[q]));
        }
        stencil_of[q] = it->second;
    }
    per_knot.clear();
    // per distinct stencil side: Psi^-1 (independent dense inversions, spread over host threads).
    // The same interface segments bound several stopping rectangles; a side whose entry
    // positions, Gaussian widths and conditioning limit equal an earlier side's bit for bit has
    // the same Psi, so its inverse (and any shape shrink) is computed once and copied.
    std::vector<std::pair<int, int>> jobs;
    std::vector<std::pair<std::pair<int, int>, int>> dups;
    {
        std::map<std::vector<uint64_t>, int> side_job;
        for (int t = 0; t < (int)stencils.size(); ++t)
            for (int d = 0; d < 4; ++d) {
                const Stencil &s = stencils[t];
                if (s.idx[d].empty()) continue;
                std::vector<uint64_t> key;
                key.reserve(3 + 3 * s.px[d].size());
                key.push_back(f64_bits(s.cshape[d]));
                key.push_back(f64_bits(s.cond_max));
                key.push_back(s.px[d].size());
                for (double v : s.px[d]) key.push_back(f64_bits(v));
                for (double v : s.py[d]) key.push_back(f64_bits(v));
                for (double v : s.cent[d]) key.push_back(f64_bits(v));
                auto it = side_job.find(key);
                if (it == side_job.end()) {
                    stencils[t].side_id[d] = (int)jobs.size();
                    side_job.emplace(std::move(key), (int)jobs.size());
                    jobs.emplace_back(t, d);
                } else {
                    stencils[t].side_id[d] = it->second;
                    dups.push_back({{t, d}, it->second});
                }
            }
    }
    std::vector<char> ok(jobs.size(), 1);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t j; (j = next.fetch_add(1)) < jobs.size();) {
            Stencil &s = stencils[jobs[j].first];
            const int d = jobs[j].second;
            const std::vector<double> *cw = s.cent[d].empty() ? nullptr : &s.cent[d];
            bool good = psi_inverse(s.px[d], s.py[d], s.cshape[d], s.ipsi[d], s.cond[d], cw);
            for (int t = 0; s.cond_max > 0.0 && t < 40 && (!good || s.cond[d] > s.cond_max); ++t) {
                s.cshape[d] *= 0.8;
                for (double &v : s.cent[d]) v *= 0.8;
                good = psi_inverse(s.px[d], s.py[d], s.cshape[d], s.ipsi[d], s.cond[d], cw);
            }
            ok[j] = good && !(s.cond_max > 0.0 && s.cond[d] > s.cond_max) ? 1 : 0;
        }
    };
    const size_t nth = std::min<size_t>(nth_all, jobs.size() / 4 + 1);
    std::vector<std::thread> pool;
    for (size_t i = 1; i < nth; ++i) pool.emplace_back(work);
    work();
    for (std::thread &th : pool) th.join();
    for (const auto &dj : dups) {
        const Stencil &src = stencils[jobs[dj.second].first];
        const int ds = jobs[dj.second].second;
        Stencil &s = stencils[dj.first.first];
        const int d = dj.first.second;
        s.cshape[d] = src.cshape[ds];
        s.cent[d] = src.cent[ds];
        s.ipsi[d] = src.ipsi[ds];
        s.cond[d] = src.cond[ds];
        if (!ok[dj.second]) { s.singular = true; any_singular = true; }
    }
    for (size_t j = 0; j < jobs.size(); ++j)
        if (!ok[j]) { stencils[jobs[j].first].singular = true; any_singular = true; }
    // per stencil: CSR row and column -> side entry maps
    for (Stencil &s : stencils) {
        std::vector<int> ent;
        for (int d = 0; d < 4; ++d) ent.insert(ent.end(), s.idx[d].begin(), s.idx[d].end());
        s.cols = ent;
        std::sort(s.cols.begin(), s.cols.end());
        s.cols.erase(std::unique(s.cols.begin(), s.cols.end()), s.cols.end());
        const int nc = (int)s.cols.size();
        s.e1.assign(nc, -1); s.e2.assign(nc, -1); s.e1_sum.assign(nc, -1); s.e2_sum.assign(nc, -1);
        for (int e = 0; e < (int)ent.size(); ++e) {
            int c = (int)(std::lower_bound(s.cols.begin(), s.cols.end(), ent[e]) - s.cols.begin());
            s.e1[c] = e;  // G_return_withrep's "=+" keeps the last side (order N,S,E,W)
            if (s.e1_sum[c] < 0) s.e1_sum[c] = e;
            else s.e2_sum[c] = e;
        }
    }
    row_ptr.assign(nknots_ + 1, 0);
    for (int q = 0; q < nknots_; ++q) {
        int nc = (on_boundary[q] || stencil_of[q] < 0) ? 0 : (int)stencils[stencil_of[q]].cols.size();
        row_ptr[q + 1] = row_ptr[q] + nc;
    }
    cols.resize(row_ptr[nknots_]);
    for (int q = 0; q < nknots_; ++q)
        if (row_ptr[q + 1] > row_ptr[q])
            std::copy(stencils[stencil_of[q]].cols.begin(), stencils[stencil_of[q]].cols.end(), cols.begin() + row_ptr[q]);
}

// Labels_Stencil (PDDAlgorithms/PDDSparseGM.cpp:1330-1433)
std::map<Dir, std::vector<std::array<int, 2>>> Mesh::labels_stencil(int k) const {
    std::map<Dir, std::vector<std::array<int, 2>>> out;
    std::vector<int> in;
    for (int i = 0; i < (int)ifs_.size(); ++i)
        if (ifs_[i].In(k)) in.push_back(i);
    auto keep = [&](Dir d, const std::array<int, 2> &l) { if (imap_.count(l)) out[d].push_back(l); };
    if (in.size() == 1) {
        const Interface &f = ifs_[in[0]];
        for (auto &l : f.N()) keep(North, l);
        for (auto &l : f.S()) keep(South, l);
        for (auto &l : f.E()) keep(East, l);
        for (auto &l : f.W()) keep(West, l);
    } else if (in.size() == 4) {
        int maxh = ifs_[in[0]].label[0], maxv = ifs_[in[0]].label[1];
        for (int t = 1; t < 4; ++t) { maxh = std::max(maxh, ifs_[in[t]].label[0]); maxv = std::max(maxv, ifs_[in[t]].label[1]); }
        int east = -1, west = -1, north = -1, south = -1;
        for (int t = 0; t < 4; ++t) {
            const auto &l = ifs_[in[t]].label;
            if (l[1] % 2 == 1) { if (l[0] == maxh) east = in[t]; else west = in[t]; }
            else { if (l[1] == maxv) north = in[t]; else south = in[t]; }
        }
        keep(North, ifs_[west].N()[0]);
        keep(North, ifs_[east].N()[0]);
        keep(South, ifs_[west].S()[0]);
        keep(South, ifs_[east].S()[0]);
        keep(East, ifs_[south].E()[0]);
        keep(East, ifs_[north].E()[0]);
        keep(West, ifs_[south].W()[0]);
        keep(West, ifs_[north].W()[0]);
    }
    return out;
}

// Compute_Stencil (PDDAlgorithms/PDDSparseGM.cpp:1845-2041) + Stencil::Init (Meshes/stencil.cpp:11-85)
Stencil Mesh::compute_stencil(int k) const {
    auto labels = labels_stencil(k);
    Stencil s;
    for (int i = 0; i < 4; ++i) s.rect[i] = glob[i];
    if (!labels[South].empty()) s.rect[1] = ifs_[find(labels[South][0][0], labels[South][0][1])].py[0];
    if (!labels[West].empty()) s.rect[0] = ifs_[find(labels[West][0][0], labels[West][0][1])].px[0];
    if (!labels[North].empty()) { const Interface &f = ifs_[find(labels[North].back()[0], labels[North].back()[1])]; s.rect[3] = f.py.back(); }
    if (!labels[East].empty()) { const Interface &f = ifs_[find(labels[East].back()[0], labels[East].back()[1])]; s.rect[2] = f.px.back(); }
    const double elong = 0.5;  // STEN_ELONG (PDDAlgorithms/PDDSparseGM.hpp:5)
    for (int d = 0; d < 4; ++d) {
        s.kind[d] = -3;
        const auto &L = labels[(Dir)d];
        if (L.empty()) continue;
        std::vector<int> ti;
        std::vector<double> tx, ty;
        auto push_range = [&](const Interface &f, int lo, int hi) {
            for (int q = lo; q < hi; ++q) { ti.push_back(f.index[q]); tx.push_back(f.px[q]); ty.push_back(f.py[q]); }
        };
        std::array<int, 2> pre = ifs_[find(L[0][0], L[0][1])].label;
        if (d <= South) pre[0] -= 1; else pre[1] -= 2;
        int ipre = find(pre[0], pre[1]);
        if (ipre >= 0) {
            s.kind[d]++;
            const Interface &f = ifs_[ipre];
            push_range(f, (int)((1 - elong) * (double)(f.index.size() - 1)), (int)f.index.size());
        }
        std::array<int, 2> last = L[0];
        for (const auto &l : L) {
            const Interface &f = ifs_[find(l[0], l[1])];
            push_range(f, 0, (int)f.index.size());
            s.kind[d] += 2;
            last = l;
        }
        std::array<int, 2> post = last;
        if (d <= South) post[0] += 1; else post[1] += 2;
        int ipost = find(post[0], post[1]);
        if (ipost >= 0) {
            const Interface &f = ifs_[ipost];
            if (d != West || !f.index.empty()) s.kind[d]++;
            push_range(f, 0, (int)(elong * (double)f.index.size()));
        }
        // Stencil::Init: index -> last written position, then std::unique (consecutive)
        std::map<int, std::pair<double, double>> pos;
        for (size_t t = 0; t < ti.size(); ++t) pos[ti[t]] = {tx[t], ty[t]};
        std::vector<int> u = ti;
        u.erase(std::unique(u.begin(), u.end()), u.end());
        s.idx[d] = u;
        for (int v : u) { s.px[d].push_back(pos[v].first); s.py[d].push_back(pos[v].second); }
    }
    return s;
}

bool psi_inverse(const std::vector<double> &px, const std::vector<double> &py, double c, std::vector<double> &ipsi,
                 double &cond, const std::vector<double> *cw) {
    const int m = (int)px.size();
    std::vector<double> A((size_t)m * m), LU;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double dx = px[i] - px[j], dy = py[i] - py[j];
            double r = std::sqrt(dx * dx + dy * dy);
            const double cj = cw ? (*cw)[j] : c;
            A[(size_t)i * m + j] = dexp(-(r * r) / (cj * cj));
        }
    LU = A;
    std::vector<int> pr(m), pc(m);
    for (int i = 0; i < m; ++i) pr[i] = pc[i] = i;
    double maxpiv = 0.0;
    bool ok = true;
    for (int kk = 0; kk < m; ++kk) {
        int bi = kk, bj = kk;
        double best = -1.0;
        for (int i = kk; i < m; ++i)
            for (int j = kk; j < m; ++j) {
                double v = std::fabs(LU[(size_t)i * m + j]);
                if (v > best) { best = v; bi = i; bj = j; }
            }
        if (kk == 0) maxpiv = best;
        if (best <= maxpiv * 2.220446049250313e-16 * m) { ok = false; break; }
        if (bi != kk) { for (int j = 0; j < m; ++j) std::swap(LU[(size_t)kk * m + j], LU[(size_t)bi * m + j]); std::swap(pr[kk], pr[bi]); }
        if (bj != kk) { for (int i = 0; i < m; ++i) std::swap(LU[(size_t)i * m + kk], LU[(size_t)i * m + bj]); std::swap(pc[kk], pc[bj]); }
        const double piv = LU[(size_t)kk * m + kk];
        for (int i = kk + 1; i < m; ++i) {
            double l = LU[(size_t)i * m + kk] / piv;
            LU[(size_t)i * m + kk] = l;
            for (int j = kk + 1; j < m; ++j) LU[(size_t)i * m + j] -= l * LU[(size_t)kk * m + j];
        }
    }
    ipsi.assign((size_t)m * m, 0.0);
    if (!ok) { cond = INFINITY; return false; }
    std::vector<double> col(m);
    for (int e = 0; e < m; ++e) {
        for (int i = 0; i < m; ++i) col[i] = (pr[i] == e) ? 1.0 : 0.0;
        for (int i = 0; i < m; ++i) { double s = col[i]; for (int j = 0; j < i; ++j) s -= LU[(size_t)i * m + j] * col[j]; col[i] = s; }
        for (int i = m - 1; i >= 0; --i) { double s = col[i]; for (int j = i + 1; j < m; ++j) s -= LU[(size_t)i * m + j] * col[j]; col[i] = s / LU[(size_t)i * m + i]; }
        for (int i = 0; i < m; ++i) ipsi[(size_t)pc[i] * m + e] = col[i];
    }
    double na = 0.0, ni = 0.0;
    for (size_t q = 0; q < A.size(); ++q) { na += A[q] * A[q]; ni += ipsi[q] * ipsi[q]; }
    cond = std::sqrt(na) * std::sqrt(ni);
    return true;
}

}  // namespace pdds
