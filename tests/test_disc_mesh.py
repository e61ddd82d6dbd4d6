"""Disc meshes (the CircleMesh layout) and annular-sector stopping domains.

The reference's Meshes/Circle*.hpp files are empty, so there is no reference geometry to match;
the layout is pinned by its invariants (counts, every side entry on its side, every knot inside
its own stencil, bounded conditioning) and the sector distance by closed forms.  A Feynman-Kac
estimate on a sector stencil with the exact u at the exit point converges to u at the knot,
which fails for a wrong projection, normal or Gobet shift.
"""
import numpy as np
import pytest

from oracle import oracle as O

R6 = 1.0 / 6.0


@pytest.fixture(scope="module")
def disc():
    from pdds_amd.engine import Mesh
    return Mesh.disc(1.0, 6, 24, 11, 8, 64, 1.0)


def test_disc_counts_and_positions(disc):
    m = disc
    r = np.hypot(m.x0, m.y0)
    assert np.all(r <= 1.0 + 1e-12)
    assert m.on_boundary.sum() == 24 and np.allclose(r[m.on_boundary], 1.0)
    # 120 arcs + 144 ray segments of 8..64 interior knots, 120 crossings, the centre, 24 rim knots
    n_int = m.n_knots - 120 - 1 - 24
    assert np.sum((m.x0 == 0) & (m.y0 == 0)) == 1
    assert 264 * 8 <= n_int <= 264 * 64
    # crossings sit on the inner circles at the ray angles
    rings = np.round(r * 6.0)
    on_circle = np.abs(r * 6.0 - rings) < 1e-9
    ang = np.arctan2(m.y0, m.x0) / (2 * np.pi / 24)
    on_ray = np.abs(ang - np.round(ang)) < 1e-9
    assert np.sum(on_circle & on_ray & ~m.on_boundary & (rings >= 1) & (rings <= 5)) == 120
    assert not m.singular


def test_disc_stencils_are_sectors_around_their_knot(disc):
    m = disc
    centre = int(np.flatnonzero((m.x0 == 0) & (m.y0 == 0))[0])
    for k in list(range(0, m.n_knots, 97)) + [m.n_knots - 25, m.n_knots - 30, centre]:
        st = m.stencil(k)
        if m.on_boundary[k]:
            assert st["msz"].sum() == 0
            continue
        g = st["domain"]
        r0, r1, ca, sa, cb, sb = g
        if r0 < 0:   # the centre knot: the whole first disc, its circle the one side
            assert m.x0[k] == 0 and m.y0[k] == 0 and st["msz"][1:].sum() == 0
            assert np.allclose(np.hypot(st["px"], st["py"]), r1, atol=1e-12)
            continue
        assert st["geom"] == 1 and r1 > r0 >= 0.0
        # the knot is strictly inside its stopping domain
        assert O.sector_dist(g, m.x0[k], m.y0[k])[0] < -1e-6
        off = np.concatenate([[0], np.cumsum(st["msz"])])
        for side in range(4):
            px, py = st["px"][off[side]:off[side + 1]], st["py"][off[side]:off[side + 1]]
            rr = np.hypot(px, py)
            if side == 0:
                assert (len(px) == 0) == (abs(r1 - 1.0) < 1e-12)
                assert np.allclose(rr, r1, atol=1e-12)
            elif side == 1:
                assert (len(px) == 0) == (r0 == 0.0)
                assert np.allclose(rr, r0, atol=1e-12)
            else:
                c, s = (cb, sb) if side == 2 else (ca, sa)
                assert np.allclose(c * py - s * px, 0.0, atol=1e-12)
                t = c * px + s * py
                assert np.all(t >= r0 - 1e-12) and np.all(t <= r1 + 1e-12)
            if len(px) > 1:   # in order along the side
                key = np.arctan2(py, px) if side <= 1 else rr
                d = np.diff(np.unwrap(key) if side <= 1 else key)
                assert np.all(d > 0)
        assert np.all(st["cond"] < 1e11)
        cols = m.row_cols(k)
        assert np.array_equal(cols, np.unique(st["idx"]))


def test_sector_dist_closed_forms():
    a, b = 0.0, 0.5                      # rays at 0 and 0.5 rad
    g = np.array([0.3, 0.6, np.cos(a), np.sin(a), np.cos(b), np.sin(b)])
    # radius 0.5, angle 0.25: distances 0.1 (outer), 0.2 (inner), 0.5 sin 0.25 = 0.124 (either
    # ray) -> the outer arc is nearest
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.5 * np.cos(0.25), 0.5 * np.sin(0.25))
    assert dd == pytest.approx(-0.1, abs=1e-15) and kd == 0
    assert (ex, ey) == pytest.approx((0.6 * np.cos(0.25), 0.6 * np.sin(0.25)), abs=1e-15)
    assert (nx, ny) == pytest.approx((np.cos(0.25), np.sin(0.25)), abs=1e-15)
    # radius 0.5, angle 0.2: ray a at 0.5 sin 0.2 = 0.0993 beats the outer arc at 0.1
    assert O.sector_dist(g, 0.5 * np.cos(0.2), 0.5 * np.sin(0.2))[5] == 3
    # near ray a: radius 0.45, angle 0.05 -> distance 0.45 sin 0.05, projection on the ray
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.45 * np.cos(0.05), 0.45 * np.sin(0.05))
    assert kd == 3 and dd == pytest.approx(-0.45 * np.sin(0.05), abs=1e-15)
    assert (ex, ey) == pytest.approx((0.45 * np.cos(0.05), 0.0), abs=1e-15) and (nx, ny) == pytest.approx((0, -1))
    # near ray b
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.45 * np.cos(0.47), 0.45 * np.sin(0.47))
    assert kd == 2 and dd == pytest.approx(-0.45 * np.sin(0.03), abs=1e-15)
    assert (nx, ny) == pytest.approx((-np.sin(b), np.cos(b)), abs=1e-15)
    # near the inner arc, and just outside it
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.31 * np.cos(0.25), 0.31 * np.sin(0.25))
    assert kd == 1 and dd == pytest.approx(-0.01, abs=1e-15) and (nx, ny) == pytest.approx((-np.cos(0.25), -np.sin(0.25)))
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.29 * np.cos(0.25), 0.29 * np.sin(0.25))
    assert kd == 1 and dd == pytest.approx(0.01, abs=1e-15)
    assert np.hypot(ex, ey) == pytest.approx(0.3, abs=1e-15)
    # outside beyond ray a, projection clamped into [r0, r1] on the ray
    dd, ex, ey, nx, ny, kd = O.sector_dist(g, 0.7, -0.05)
    assert dd > 0 and (ex, ey) == pytest.approx((0.6, 0.0), abs=1e-15)
    # a centre sector (r0 = 0) has no inner side
    g0 = np.array([0.0, 0.6, np.cos(a), np.sin(a), np.cos(b), np.sin(b)])
    assert O.sector_dist(g0, 0.01 * np.cos(0.25), 0.01 * np.sin(0.25))[5] in (2, 3)
    # the whole disc (r0 < 0): only its circle, whatever the angle
    gw = np.array([-1.0, 0.5, 1.0, 0.0, 1.0, 0.0])
    for t in (0.0, 2.0, -2.5):
        dd, ex, ey, nx, ny, kd = O.sector_dist(gw, 0.2 * np.cos(t), 0.2 * np.sin(t))
        assert kd == 0 and dd == pytest.approx(-0.3, abs=1e-15)
        assert (ex, ey) == pytest.approx((0.5 * np.cos(t), 0.5 * np.sin(t)), abs=1e-15)


def test_feynman_kac_on_a_sector_converges_to_u(disc):
    """Z + xi + Y u(X_tau) with the exact u at the exit estimates u(knot): the sector distance,
    its projection and normal and the Gobet shift together must be right (weak order h)."""
    m = disc
    pb = O.problem(O.PDE_SINSIN)
    for k in (5, 2000, 7000):
        st = m.stencil(k)
        errs = []
        for h in (4e-4, 1e-4):
            p = O.simulate(pb, O.run_params(h, 3, threads=8), m.x0[k], m.y0[k], st["domain"], k, 0, 20000, geom=1)
            u = np.array([O.pde_eval(pb, a, b)["u"] for a, b in zip(p["xt"], p["yt"])])
            est = p["Z"] + p["xi"] + p["Y"] * u
            se = est.std() / np.sqrt(len(est))
            errs.append((est.mean() - O.pde_eval(pb, m.x0[k], m.y0[k])["u"], se))
        # the estimate approaches u at weak order h: the bias at h = 1e-4 is well below the one
        # at 4e-4 (ratio near 4, at least 2 here) and small
        assert abs(errs[1][0]) < 0.5 * abs(errs[0][0]) + 3 * errs[1][1], (k, errs)
        assert abs(errs[1][0]) < 5e-3, (k, errs)
        # exits land on the stopping domain's boundary
        d = np.array([O.sector_dist(st["domain"], a, b)[0] for a, b in zip(p["xt"][:500], p["yt"][:500])])
        assert np.all(np.abs(d) < 1e-12)


def test_disc_side_interpolation_is_accurate(disc):
    """Each side's RBF interpolant of a smooth u (the cardinal functions H of the G rows) is
    accurate between its end knots, including sides that join interfaces of very different knot
    counts (per-centre widths; a single width per side left errors near 0.2 there).  A ray side
    from the centre has no knot at the centre itself, so it is checked from its first knot."""
    m = disc
    u = lambda x, y: np.exp(x) * np.cos(y)   # noqa: E731
    worst = 0.0
    for k in range(0, m.n_knots, 41):
        if m.on_boundary[k]:
            continue
        st = m.stencil(k)
        g = st["domain"]
        off = np.concatenate([[0], np.cumsum(st["msz"])])
        for s in range(4):
            sl = slice(off[s], off[s + 1])
            px, py, c = st["px"][sl], st["py"][sl], st["cent"][sl]
            if len(px) < 2:
                continue
            if s <= 1:
                r = g[1] if s == 0 else g[0]
                t0, t1 = np.arctan2(g[3], g[2]), np.arctan2(g[5], g[4])
                t1 = t1 + 2 * np.pi if t1 < t0 else t1
                t = np.linspace(t0, t1, 300)
                mx, my = r * np.cos(t), r * np.sin(t)
            else:
                ca, sa = (g[4], g[5]) if s == 2 else (g[2], g[3])
                rr = np.linspace(np.hypot(px[0], py[0]), g[1], 300)
                mx, my = rr * ca, rr * sa
            phi = np.exp(-((px[None, :] - mx[:, None]) ** 2 + (py[None, :] - my[:, None]) ** 2) / c[None, :] ** 2)
            interp = (phi @ st["ipsi"][s]) @ u(px, py)
            worst = max(worst, np.abs(interp - u(mx, my)).max())
    assert worst < 1e-2, worst
