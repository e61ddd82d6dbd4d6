// disc_mesh.cpp -- PDDSparse geometry on a disc (the CircleMesh layout; the reference's
// Meshes/Circle*.hpp files are empty, so the layout is this engine's own):
//   - subdomains: rings rho_i <= r <= rho_{i+1} (rho_i = R i / n_rings) times sectors
//     t_k <= theta <= t_{k+1} (t_k = 2 pi k / n_sectors);
//   - interfaces: the arcs of the inner circles between neighbouring rays, and the ray
//     segments between neighbouring circles (the centre is no knot);
//   - knots: n_I equispaced interior knots per interface (n_I uniform in [kmin, kmax] from a
//     seeded hash), a knot at every crossing of an inner circle with a ray, one at the centre
//     (where the rays meet), and a global-boundary knot where each ray meets the outer circle;
//   - stencil of a knot: the union of the subdomains around its interface (an annular sector
//     spanning two rings or two sectors, or both for a crossing), with four sides: outer arc (0),
//     inner arc (1), end ray b (2) and start ray a (3), each listing the knots on it in order.
//     An outer arc on the disc boundary is a global side; a ring-0 stencil has no inner side and
//     its rays start at the centre knot.  The centre knot's stencil is the whole first disc,
//     whose circle is its one side.
#include "mesh.h"

#include <cmath>

namespace pdds {

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

}  // namespace

Mesh Mesh::disc(double R, int nr, int ns, uint32_t seed, int kmin, int kmax, double shape) {
    Mesh m;
    m.geom = GeomDisc;
    m.radius = R;
    m.glob[0] = -R; m.glob[1] = -R; m.glob[2] = R; m.glob[3] = R;
    for (int i = 0; i < 4; ++i) m.fac[i] = shape;
    const double PI = 3.14159265358979323846;
    std::vector<double> rho(nr + 1), ca(ns), sa(ns);
    for (int i = 0; i <= nr; ++i) rho[i] = i == nr ? R : R * (double)i / (double)nr;
    for (int k = 0; k < ns; ++k) { ca[k] = std::cos(2.0 * PI * k / ns); sa[k] = std::sin(2.0 * PI * k / ns); }
    auto count = [&](uint64_t iface) {
        return kmin + (int)(splitmix64(((uint64_t)seed << 32) ^ iface) % (uint64_t)(kmax - kmin + 1));
    };
    auto wrap = [&](int k) { return ((k % ns) + ns) % ns; };
    // knots: arc interiors, ray interiors, crossings, boundary knots
    std::vector<std::vector<int>> arc((size_t)(nr + 1) * ns), ray((size_t)nr * ns);
    std::vector<int> cross((size_t)(nr + 1) * ns, -1), bnd(ns, -1);
    std::vector<double> X, Y;
    std::vector<int> onb;
    auto add = [&](double x, double y, int b) { X.push_back(x); Y.push_back(y); onb.push_back(b); return (int)X.size() - 1; };
    uint64_t iface = 0;
    for (int i = 1; i < nr; ++i)
        for (int k = 0; k < ns; ++k, ++iface) {
            const int n = count(iface);
            for (int j = 0; j < n; ++j) {
                const double t = 2.0 * PI * ((double)k + (double)(j + 1) / (double)(n + 1)) / ns;
                arc[(size_t)i * ns + k].push_back(add(rho[i] * std::cos(t), rho[i] * std::sin(t), 0));
            }
        }
    for (int i = 0; i < nr; ++i)
        for (int k = 0; k < ns; ++k, ++iface) {
            const int n = count(iface);
            for (int j = 0; j < n; ++j) {
                const double r = rho[i] + (rho[i + 1] - rho[i]) * (double)(j + 1) / (double)(n + 1);
                ray[(size_t)i * ns + k].push_back(add(r * ca[k], r * sa[k], 0));
            }
        }
    for (int i = 1; i < nr; ++i)
        for (int k = 0; k < ns; ++k) cross[(size_t)i * ns + k] = add(rho[i] * ca[k], rho[i] * sa[k], 0);
    const int centre = add(0.0, 0.0, 0);
    for (int k = 0; k < ns; ++k) bnd[k] = add(R * ca[k], R * sa[k], 1);
    m.nknots_ = (int)X.size();
    m.x0 = X; m.y0 = Y; m.npx = X; m.npy = Y;
    m.on_boundary = onb;
    // side lists: the knots on circle i between rays k0 and k0 + span, and on ray k between
    // circles i0 and i1, in order, endpoints included
    auto arc_side = [&](int i, int k0, int span) {
        std::vector<int> v;
        for (int s = 0; s < span; ++s) {
            const int k = wrap(k0 + s);
            v.push_back(cross[(size_t)i * ns + k]);
            const auto &a = arc[(size_t)i * ns + k];
            v.insert(v.end(), a.begin(), a.end());
        }
        v.push_back(cross[(size_t)i * ns + wrap(k0 + span)]);
        return v;
    };
    auto ray_side = [&](int k, int i0, int i1) {
        std::vector<int> v;
        for (int i = i0; i < i1; ++i) {
            v.push_back(i >= 1 ? cross[(size_t)i * ns + k] : centre);
            const auto &a = ray[(size_t)i * ns + k];
            v.insert(v.end(), a.begin(), a.end());
        }
        v.push_back(i1 == nr ? bnd[k] : cross[(size_t)i1 * ns + k]);
        return v;
    };
    std::vector<Stencil> per_knot(m.nknots_);
    std::vector<char> has(m.nknots_, 0);
    auto make = [&](int q, int i0, int i1, int ka, int span) {
        const int kb = wrap(ka + span);
        Stencil s;
        s.geo[0] = rho[i0]; s.geo[1] = rho[i1];
        s.geo[2] = ca[ka]; s.geo[3] = sa[ka]; s.geo[4] = ca[kb]; s.geo[5] = sa[kb];
        std::vector<int> side[4];
        if (i1 < nr) side[0] = arc_side(i1, ka, span);
        else s.gmask |= 1;
        if (i0 > 0) side[1] = arc_side(i0, ka, span);
        side[2] = ray_side(kb, i0, i1);
        side[3] = ray_side(ka, i0, i1);
        const double dt = 2.0 * PI * span / ns;
        const double len[4] = {rho[i1] * dt, rho[i0] * dt, rho[i1] - rho[i0], rho[i1] - rho[i0]};
        for (int d = 0; d < 4; ++d) {
            s.kind[d] = 0;
            s.idx[d] = side[d];
            for (int v : side[d]) { s.px[d].push_back(X[v]); s.py[d].push_back(Y[v]); }
            // per-centre widths from the local spacing along the side: the sides join interfaces
            // of different knot counts, and one width for all would leave the sparse part
            // under-resolved or the dense part singular
            const size_t n = side[d].size();
            s.cshape[d] = n > 1 ? shape * 3.5 * len[d] / (double)(n - 1) : 1.0;
            if (n > 1) {
                std::vector<double> gap(n - 1);
                for (size_t j = 0; j + 1 < n; ++j) gap[j] = std::hypot(s.px[d][j + 1] - s.px[d][j], s.py[d][j + 1] - s.py[d][j]);
                s.cent[d].resize(n);
                for (size_t j = 0; j < n; ++j) {
                    const double loc = j == 0 ? gap[0] : (j + 1 == n ? gap[n - 2] : 0.5 * (gap[j - 1] + gap[j]));
                    s.cent[d][j] = shape * 3.5 * loc;
                }
            }
        }
        s.cond_max = 1e11;
        // bounding box: corners, plus the outer circle where it crosses an axis inside the sector
        double bx0 = 1e300, by0 = 1e300, bx1 = -1e300, by1 = -1e300;
        auto inc = [&](double x, double y) { bx0 = std::min(bx0, x); by0 = std::min(by0, y); bx1 = std::max(bx1, x); by1 = std::max(by1, y); };
        for (double r : {rho[i0], rho[i1]}) { inc(r * ca[ka], r * sa[ka]); inc(r * ca[kb], r * sa[kb]); }
        for (int ax = 0; ax < 4; ++ax) {
            const double t = ax * ns / 4.0;   // axis angle in sector units
            double rel = t - ka;
            while (rel < 0) rel += ns;
            if (rel <= span) inc(rho[i1] * std::cos(ax * PI / 2), rho[i1] * std::sin(ax * PI / 2));
        }
        for (int i = 0; i < 4; ++i) s.rect[i] = 0.0;
        s.bbox[0] = bx0; s.bbox[1] = by0; s.bbox[2] = bx1; s.bbox[3] = by1;
        per_knot[q] = std::move(s);
        has[q] = 1;
    };
    for (int i = 1; i < nr; ++i)
        for (int k = 0; k < ns; ++k) {
            for (int q : arc[(size_t)i * ns + k]) make(q, i - 1, i + 1, k, 1);
            make(cross[(size_t)i * ns + k], i - 1, i + 1, wrap(k - 1), 2);
        }
    for (int i = 0; i < nr; ++i)
        for (int k = 0; k < ns; ++k)
            for (int q : ray[(size_t)i * ns + k]) make(q, i, i + 1, wrap(k - 1), 2);
    {
        // the centre: the whole first disc, its circle (every crossing and arc knot on it, in
        // angle order) the one side
        Stencil s;
        s.geo[0] = -1.0; s.geo[1] = rho[1]; s.geo[2] = 1.0; s.geo[3] = 0.0; s.geo[4] = 1.0; s.geo[5] = 0.0;
        std::vector<int> side;
        for (int k = 0; k < ns; ++k) {
            side.push_back(cross[(size_t)ns + k]);
            const auto &a = arc[(size_t)ns + k];
            side.insert(side.end(), a.begin(), a.end());
        }
        const size_t n = side.size();
        for (int d = 0; d < 4; ++d) s.kind[d] = 0;
        s.idx[0] = side;
        for (int v : side) { s.px[0].push_back(X[v]); s.py[0].push_back(Y[v]); }
        s.cshape[0] = shape * 3.5 * (2.0 * PI * rho[1]) / (double)n;
        s.cent[0].resize(n);
        for (size_t j = 0; j < n; ++j) {
            const size_t jp = (j + 1) % n, jm = (j + n - 1) % n;
            const double g1 = std::hypot(s.px[0][jp] - s.px[0][j], s.py[0][jp] - s.py[0][j]);
            const double g0 = std::hypot(s.px[0][j] - s.px[0][jm], s.py[0][j] - s.py[0][jm]);
            s.cent[0][j] = shape * 3.5 * 0.5 * (g0 + g1);
        }
        s.cond_max = 1e11;
        for (int i = 0; i < 4; ++i) s.rect[i] = 0.0;
        s.bbox[0] = -rho[1]; s.bbox[1] = -rho[1]; s.bbox[2] = rho[1]; s.bbox[3] = rho[1];
        per_knot[centre] = std::move(s);
        has[centre] = 1;
    }
    m.finalize(per_knot, has);
    return m;
}

}  // namespace pdds
