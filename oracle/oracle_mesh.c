/*
 * oracle_mesh.c -- TEST INFRASTRUCTURE (the oracle). Not part of the product.
 *
 * Restatement of the reference's interface/knot/stencil bookkeeping for the
 * configuration used by main_EXTENDED (INTERSECTIONS_YES, no HAIRY, no
 * SQUARE_PATCHES, STEN_ELONG = 0.5; PDDAlgorithms/PDDSparseGM.hpp:4-75):
 *   interfaces, knot numbering, positions   PDDAlgorithms/PDDSparseGM.cpp:346-689
 *   knot start positions (dispatch order)   PDDAlgorithms/PDDSparseGM.cpp:705-756
 *   Node_Position                           PDDAlgorithms/PDDSparseGM.cpp:3156-3167
 *   Interface::N/S/E/W                      Meshes/interface.cpp:16-93
 *   Get_Interfaces / Labels_Stencil         PDDAlgorithms/PDDSparseGM.cpp:1316-1433
 *   Compute_Stencil                         PDDAlgorithms/PDDSparseGM.cpp:1845-2041
 *   Stencil::Init (consecutive dedup)       Meshes/stencil.cpp:11-85
 * Square grids only (N_subdo and N_node set both directions,
 * PDDAlgorithms/PDDSparseGM.cpp:303-306, 326-328).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct {
    int a, b;          /* 2-D label */
    int n;             /* knots on the interface */
    int *idx;
    double *px, *py;
} orc_iface;

struct orc_mesh {
    int n_sub, n_node;
    double sw[2], ne[2];
    int nif;
    orc_iface *ifs;
    int nknots;
    double *x0, *y0, *npx, *npy;
    int *onb;
};

static int orc_find(const orc_mesh *m, int a, int b) {
    for (int i = 0; i < m->nif; ++i) if (m->ifs[i].a == a && m->ifs[i].b == b) return i;
    return -1;
}

orc_mesh *orc_mesh_create(double sw_x, double sw_y, double ne_x, double ne_y, int n_sub, int n_node) {
    if (n_sub < 2 || n_node < 2) return NULL;
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    int n = n_sub, k = n_node;
    m->n_sub = n; m->n_node = k;
    m->sw[0] = sw_x; m->sw[1] = sw_y; m->ne[0] = ne_x; m->ne[1] = ne_y;
    m->nif = n * (n - 1) + n * (n - 1);
    m->ifs = (orc_iface *)calloc((size_t)m->nif, sizeof(orc_iface));
    double *stx = (double *)malloc(sizeof(double) * m->nif), *sty = (double *)malloc(sizeof(double) * m->nif);
    double *enx = (double *)malloc(sizeof(double) * m->nif), *eny = (double *)malloc(sizeof(double) * m->nif);
    double L0x = (ne_x - sw_x) / (1.0 * n), L1y = (ne_y - sw_y) / (1.0 * n);
    double P0x = sw_x + 0.0, P0y = sw_y + L1y;
    double P1x = sw_x + L0x, P1y = sw_y + 0.0;
    int cen = 0;
    for (int i = 0; i < n - 1; ++i) {
        for (int j = 0; j < n; ++j) {
            stx[cen] = P0x + L0x * j; sty[cen] = P0y + 0.0 * j;
            enx[cen] = P0x + L0x * (j + 1); eny[cen] = P0y + 0.0 * (j + 1);
            m->ifs[cen].a = j; m->ifs[cen].b = 2 * i + 1;
            cen++;
            stx[cen] = P1x + 0.0 * j; sty[cen] = P1y + L1y * j;
            enx[cen] = P1x + 0.0 * (j + 1); eny[cen] = P1y + L1y * (j + 1);
            m->ifs[cen].a = i; m->ifs[cen].b = 2 * j;
            cen++;
        }
        P0x = P0x + 0.0; P0y = P0y + L1y;
        P1x = P1x + L0x; P1y = P1y + 0.0;
    }
    for (int i = 0; i < m->nif; ++i) {
        m->ifs[i].idx = (int *)malloc(sizeof(int) * (size_t)(k + 2));
        m->ifs[i].n = 0;
    }
    /* knot numbering with shared crossings */
    int *cross = (int *)malloc(sizeof(int) * (size_t)((n - 1) * (n - 1) + 1));
    int cent = 0, cross_cent = 0;
    for (int i = 0; i < n - 1; ++i) {
        for (int j = 0; j <= 2 * n - 2; j += 2) {
            orc_iface *f = &m->ifs[orc_find(m, i, j)];
            f->idx[f->n++] = cent;
            if (j != 0) cross[(j / 2 - 1) * (n - 1) + i] = cent;
            for (int q = 0; q < k - 1; ++q) { cent++; f->idx[f->n++] = cent; }
            if (j == 2 * n - 2) cent++;
        }
    }
    for (int j = 1; j <= 2 * n - 3; j += 2) {
        for (int i = 0; i < n; ++i) {
            orc_iface *f = &m->ifs[orc_find(m, i, j)];
            if (i != 0) { f->idx[f->n++] = cross[cross_cent]; cross_cent++; }
            else { f->idx[f->n++] = cent; cent++; }
            for (int q = 1; q < k - 1; ++q) { f->idx[f->n++] = cent; cent++; }
            if (i != n - 1) f->idx[f->n++] = cross[cross_cent];
            else { f->idx[f->n++] = cent; cent++; }
        }
    }
    m->nknots = cent;
    /* positions: start + accumulated increments */
    for (int i = 0; i < m->nif; ++i) {
        orc_iface *f = &m->ifs[i];
        f->px = (double *)malloc(sizeof(double) * (size_t)f->n);
        f->py = (double *)malloc(sizeof(double) * (size_t)f->n);
        double ix = (enx[i] - stx[i]) / (double)(f->n - 1), iy = (eny[i] - sty[i]) / (double)(f->n - 1);
        double vx = stx[i], vy = sty[i];
        for (int q = 0; q < f->n; ++q) { f->px[q] = vx; f->py[q] = vy; vx = vx + ix; vy = vy + iy; }
    }
    free(stx); free(sty); free(enx); free(eny); free(cross);
    /* start positions in dispatch order; Node_Position = last interface in order */
    m->x0 = (double *)calloc((size_t)m->nknots, sizeof(double));
    m->y0 = (double *)calloc((size_t)m->nknots, sizeof(double));
    m->npx = (double *)calloc((size_t)m->nknots, sizeof(double));
    m->npy = (double *)calloc((size_t)m->nknots, sizeof(double));
    m->onb = (int *)calloc((size_t)m->nknots, sizeof(int));
    for (int Iy = 0; Iy <= 2 * n - 2; ++Iy) {
        if (Iy % 2 == 0) {
            for (int Ix = 0; Ix <= n - 2; ++Ix) {
                orc_iface *f = &m->ifs[orc_find(m, Ix, Iy)];
                int lo = (Iy == 0) ? 0 : 1, hi = f->n;
                for (int q = lo; q < hi; ++q) { m->x0[f->idx[q]] = f->px[q]; m->y0[f->idx[q]] = f->py[q]; }
            }
        } else {
            for (int Ix = 0; Ix <= n - 1; ++Ix) {
                orc_iface *f = &m->ifs[orc_find(m, Ix, Iy)];
                int lo, hi;
                if (Ix == 0) { lo = 0; hi = f->n - 1; }
                else if (Ix == n - 1) { lo = 1; hi = f->n; }
                else { lo = 1; hi = f->n - 1; }
                for (int q = lo; q < hi; ++q) { m->x0[f->idx[q]] = f->px[q]; m->y0[f->idx[q]] = f->py[q]; }
            }
        }
    }
    for (int i = 0; i < m->nif; ++i) {
        orc_iface *f = &m->ifs[i];
        for (int q = 0; q < f->n; ++q) { m->npx[f->idx[q]] = f->px[q]; m->npy[f->idx[q]] = f->py[q]; }
    }
    double glob[4] = {sw_x, sw_y, ne_x, ne_y};
    for (int q = 0; q < m->nknots; ++q) {
        double ex, ey, nx, ny;
        m->onb[q] = fabs(orc_rect_dist(glob, m->x0[q], m->y0[q], &ex, &ey, &nx, &ny)) <= 1e-8;
    }
    return m;
}

void orc_mesh_destroy(orc_mesh *m) {
    if (!m) return;
    for (int i = 0; i < m->nif; ++i) { free(m->ifs[i].idx); free(m->ifs[i].px); free(m->ifs[i].py); }
    free(m->ifs); free(m->x0); free(m->y0); free(m->npx); free(m->npy); free(m->onb);
    free(m);
}

int orc_mesh_n_knots(const orc_mesh *m) { return m->nknots; }
int orc_mesh_n_interfaces(const orc_mesh *m) { return m->nif; }

void orc_mesh_knots(const orc_mesh *m, double *x0, double *y0, double *nx, double *ny, int *onb) {
    for (int q = 0; q < m->nknots; ++q) {
        if (x0) x0[q] = m->x0[q];
        if (y0) y0[q] = m->y0[q];
        if (nx) nx[q] = m->npx[q];
        if (ny) ny[q] = m->npy[q];
        if (onb) onb[q] = m->onb[q];
    }
}

/* Interface::N/S/E/W (Meshes/interface.cpp:16-93): up to two neighbour labels */
static int orc_nb(int a, int b, int dir, int out[2][2]) {
    int vert = (b % 2 == 0);
    switch (dir) {
        case 0: /* N */
            if (vert) { out[0][0] = a; out[0][1] = b + 1; out[1][0] = a + 1; out[1][1] = b + 1; return 2; }
            out[0][0] = a; out[0][1] = b + 2; return 1;
        case 1: /* S */
            if (vert) { out[0][0] = a; out[0][1] = b - 1; out[1][0] = a + 1; out[1][1] = b - 1; return 2; }
            out[0][0] = a; out[0][1] = b - 2; return 1;
        case 2: /* E */
            if (vert) { out[0][0] = a + 1; out[0][1] = b; return 1; }
            out[0][0] = a; out[0][1] = b - 1; out[1][0] = a; out[1][1] = b + 1; return 2;
        default: /* W */
            if (vert) { out[0][0] = a - 1; out[0][1] = b; return 1; }
            out[0][0] = a - 1; out[0][1] = b - 1; out[1][0] = a - 1; out[1][1] = b + 1; return 2;
    }
}

/* Full Labels_Stencil + Compute_Stencil + Stencil::Init for knot k. */
int orc_mesh_stencil(const orc_mesh *m, int k, double rect[4], int msz[4], int kind[4],
                     int *idx_out, double *px_out, double *py_out) {
    int ifl[8], ni = 0;
    for (int i = 0; i < m->nif && ni < 8; ++i)
        for (int q = 0; q < m->ifs[i].n; ++q) if (m->ifs[i].idx[q] == k) { ifl[ni++] = i; break; }
    int L[4][2][2];   /* labels per direction */
    int nL[4] = {0, 0, 0, 0};
    if (ni == 1) {
        const orc_iface *f = &m->ifs[ifl[0]];
        for (int d = 0; d < 4; ++d) {
            int out[2][2];
            int c = orc_nb(f->a, f->b, d, out);
            for (int t = 0; t < c; ++t)
                if (orc_find(m, out[t][0], out[t][1]) >= 0) {
                    L[d][nL[d]][0] = out[t][0]; L[d][nL[d]][1] = out[t][1]; nL[d]++;
                }
        }
    } else if (ni == 4) {
        int maxh = -1000000, maxv = -1000000;
        for (int t = 0; t < 4; ++t) {
            if (m->ifs[ifl[t]].a > maxh) maxh = m->ifs[ifl[t]].a;
            if (m->ifs[ifl[t]].b > maxv) maxv = m->ifs[ifl[t]].b;
        }
        int east = -1, west = -1, north = -1, south = -1;
        for (int t = 0; t < 4; ++t) {
            const orc_iface *f = &m->ifs[ifl[t]];
            if (f->b % 2 == 1) { if (f->a == maxh) east = ifl[t]; else west = ifl[t]; }
            else { if (f->b == maxv) north = ifl[t]; else south = ifl[t]; }
        }
        struct { int from; int dir; int dest; } plan[8] = {
            {west, 0, 0}, {east, 0, 0}, {west, 1, 1}, {east, 1, 1},
            {south, 2, 2}, {north, 2, 2}, {south, 3, 3}, {north, 3, 3}};
        for (int t = 0; t < 8; ++t) {
            if (plan[t].from < 0) continue;
            const orc_iface *f = &m->ifs[plan[t].from];
            int out[2][2];
            orc_nb(f->a, f->b, plan[t].dir, out);
            if (orc_find(m, out[0][0], out[0][1]) >= 0) {
                int d = plan[t].dest;
                L[d][nL[d]][0] = out[0][0]; L[d][nL[d]][1] = out[0][1]; nL[d]++;
            }
        }
    }
    /* stencil rectangle */
    rect[0] = m->sw[0]; rect[1] = m->sw[1]; rect[2] = m->ne[0]; rect[3] = m->ne[1];
    if (nL[1] > 0) rect[1] = m->ifs[orc_find(m, L[1][0][0], L[1][0][1])].py[0];
    if (nL[3] > 0) rect[0] = m->ifs[orc_find(m, L[3][0][0], L[3][0][1])].px[0];
    if (nL[0] > 0) { const orc_iface *f = &m->ifs[orc_find(m, L[0][nL[0] - 1][0], L[0][nL[0] - 1][1])]; rect[3] = f->py[f->n - 1]; }
    if (nL[2] > 0) { const orc_iface *f = &m->ifs[orc_find(m, L[2][nL[2] - 1][0], L[2][nL[2] - 1][1])]; rect[2] = f->px[f->n - 1]; }
    /* side lists with elongation */
    int total = 0;
    for (int d = 0; d < 4; ++d) {
        kind[d] = -3;
        msz[d] = 0;
        if (nL[d] == 0) continue;
        int cap = 4 * (m->n_node + 2) + 8;
        int *ti = (int *)malloc(sizeof(int) * cap);
        double *tx = (double *)malloc(sizeof(double) * cap), *ty = (double *)malloc(sizeof(double) * cap);
        int nt = 0;
        const orc_iface *first = &m->ifs[orc_find(m, L[d][0][0], L[d][0][1])];
        int pa = first->a, pb = first->b;
        if (d <= 1) pa -= 1; else pb -= 2;
        int pre = orc_find(m, pa, pb);
        if (pre >= 0) {
            kind[d]++;
            const orc_iface *f = &m->ifs[pre];
            for (int q = (int)((1 - 0.5) * (f->n - 1)); q < f->n; ++q) { ti[nt] = f->idx[q]; tx[nt] = f->px[q]; ty[nt] = f->py[q]; nt++; }
        }
        const orc_iface *last = first;
        for (int t = 0; t < nL[d]; ++t) {
            const orc_iface *f = &m->ifs[orc_find(m, L[d][t][0], L[d][t][1])];
            for (int q = 0; q < f->n; ++q) { ti[nt] = f->idx[q]; tx[nt] = f->px[q]; ty[nt] = f->py[q]; nt++; }
            kind[d] += 2;
            last = f;
        }
        int qa = last->a, qb = last->b;
        if (d <= 1) qa += 1; else qb += 2;
        int post = orc_find(m, qa, qb);
        if (post >= 0) {
            const orc_iface *f = &m->ifs[post];
            if (d != 3 || f->n > 0) kind[d]++;
            for (int q = 0; q < (int)(0.5 * f->n); ++q) { ti[nt] = f->idx[q]; tx[nt] = f->px[q]; ty[nt] = f->py[q]; nt++; }
        }
        /* Stencil::Init: map index -> last written position; std::unique (consecutive) */
        int nu = 0;
        int *ui = (int *)malloc(sizeof(int) * (nt + 1));
        for (int t = 0; t < nt; ++t) if (nu == 0 || ui[nu - 1] != ti[t]) ui[nu++] = ti[t];
        for (int t = 0; t < nu; ++t) {
            double ux = 0.0, uy = 0.0;
            for (int s = 0; s < nt; ++s) if (ti[s] == ui[t]) { ux = tx[s]; uy = ty[s]; }
            if (idx_out) { idx_out[total + t] = ui[t]; px_out[total + t] = ux; py_out[total + t] = uy; }
        }
        msz[d] = nu;
        total += nu;
        free(ti); free(tx); free(ty); free(ui);
    }
    return total;
}
