"""The knot-equation statistics of the batching form, pinned by brute force.

For reference interface values w, each path of knot k gives one sample of the knot's equation
r_p = score_p - sum_j g_j(p) w_j, with g_j(p) = -Y_p H_j(X_p) for an exit through a stencil side
(H = Psi^-T phi, the per-exit loop of Stencil::G_update, Meshes/stencil.cpp:244-307) and 0 for a
global exit.  The oracle's batched reduction forms the same quantity from sub-batch sums; its mean
must equal the per-path mean to rounding and its batch-means variance the per-path variance to
the sampling noise of nb batches (about sqrt(2 / nb)).  A wrong sign on the G term, a missing
1/n_b or a wrong column weight fails one of the two.
"""
import numpy as np
import pytest

from conftest import FIX_FAC, FIX_H, FIX_NE, FIX_NNODE, FIX_NSUB, FIX_SW
from oracle import oracle as O


def _ipsi(st):
    ip, cs = [], []
    off = 0
    for s in range(4):
        m = int(st["msz"][s])
        c = FIX_FAC[min(max(int(st["kind"][s]), 0), 3)]
        cs.append(c)
        if m == 0:
            ip.append(np.zeros((0, 0)))
            continue
        status, inv, _ = O.stencil_ipsi(st["px"][off:off + m], st["py"][off:off + m], c)
        assert status == 0
        ip.append(inv)
        off += m
    return ip, cs


def _per_path(pb, paths, st, glob, cs, ip, w_of_knot):
    """r_p for every path, from the per-exit definition (sum mode at shared corners)."""
    msz = st["msz"]
    off = np.concatenate([[0], np.cumsum(msz)])
    rect = st["rect"]
    r = np.zeros(len(paths["Y"]))
    for n in range(len(r)):
        x, y, Y = paths["xt"][n], paths["yt"][n], paths["Y"][n]
        if O.rect_dist(glob, x, y)[0] >= -1e-8:
            r[n] = paths["Z"][n] + Y * O.pde_eval(pb, x, y)["g"] + paths["xi"][n]
            continue
        side = -1
        for s, hit in ((1, abs(y - rect[1]) < 1e-8), (0, abs(y - rect[3]) < 1e-8),
                       (3, abs(x - rect[0]) < 1e-8), (2, abs(x - rect[2]) < 1e-8)):
            if hit:
                side = s
                break
        assert side >= 0
        sl = slice(off[side], off[side + 1])
        phi = np.exp(-((st["px"][sl] - x) ** 2 + (st["py"][sl] - y) ** 2) / cs[side] ** 2)
        H = ip[side].T @ phi
        g = -Y * H
        r[n] = paths["Z"][n] + paths["xi"][n] - np.sum(g * w_of_knot[st["idx"][sl]])
    return r


@pytest.mark.parametrize("knot", [40, 100])
def test_equation_mean_and_variance_match_per_path_values(knot):
    mesh = O.Mesh(FIX_SW, FIX_NE, FIX_NSUB, FIX_NNODE)
    x0, y0, px, py, onb = mesh.knots()
    pb = O.problem(O.PDE_MONEGROS)
    u = np.array([O.pde_eval(pb, a, b)["u"] for a, b in zip(px, py)])
    w = u + 0.05 * np.sin(np.arange(len(u)))     # any reference values will do
    st = mesh.stencil(knot)
    ip, cs = _ipsi(st)
    n = 20000
    paths = O.simulate(pb, O.run_params(FIX_H, 5, threads=8), x0[knot], y0[knot], st["rect"], knot, 0, n)
    got = O.reduce_batched_eq(pb, paths, st, mesh.glob, cs, ip, 32, 640, w, quirk=0)
    r = _per_path(pb, paths, st, mesh.glob, cs, ip, w)
    assert got["eq_mean"] == pytest.approx(r.mean(), rel=1e-9, abs=1e-9)
    ratio = got["eq_var"] / r.var(ddof=1)
    assert 0.75 < ratio < 1.33, ratio
    # the knot equation is far less noisy than its B entry alone: the stencil terms cancel the
    # variance the exit position brings into the score
    assert got["eq_var"] < (got["BB"] - got["B"] ** 2)
