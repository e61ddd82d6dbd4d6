"""Distance-based adaptive h (PDDAlgorithms/PDDSparseGM.cpp:4025-4029, :4386-4389) on the GPU.

With run.adaptive_h a knot at signed distance d <= -1e-4 from its stencil's stopping domain steps
with h_k = min(h, (d / (2 |beta| |N^T sigma(X0)|))^2), N the nearest side's outward normal, so
that the Gobet shift is at most half the distance.  The per-knot steps equal that rule evaluated
from the oracle's distance and coefficients; a knot's G row, variances, B and BB with its h_k equal
the oracle's paths at h_k, bit for bit, in the tiled, persistent and plain kernels; and an
assembly's rows equal the knots solved alone.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

H0 = 4e-3
SW, NE = (-1.0, -1.0), (1.0, 1.0)


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def setup(eng):
    from pdds_amd.engine import Mesh, bvp
    from pdds_amd.workloads import relative_fac, varcoef_fields
    fac = relative_fac(SW, NE, 4, 8, 1.0)
    m = Mesh(SW, NE, 4, 8, fac)
    n = 96
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    samples = varcoef_fields(X, Y)
    prob = bvp("varcoef_lut", lut=eng.lut(samples, (-1, -1, 1, 1), mode=0))
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=O.lut_dict(samples, (-1, -1, 1, 1)))
    return m, fac, prob, pb


def _rule(m, pb, beta=-0.5826):
    h = np.zeros(m.n_knots)
    for q in range(m.n_knots):
        if m.on_boundary[q]:
            continue
        d, _, _, nx, ny = O.rect_dist(m.stencil(q)["rect"], m.x0[q], m.y0[q])
        if not d <= -1e-4:
            h[q] = H0
            continue
        s = O.pde_eval(pb, m.x0[q], m.y0[q])["sigma"]
        ga, gb = nx * s[0] + ny * s[2], nx * s[1] + ny * s[3]
        r = d / (np.sqrt(ga * ga + gb * gb) * 2.0 * abs(beta))
        h[q] = min(H0, r * r)
    return h


def test_knot_steps_follow_the_distance_rule(eng, setup):
    from pdds_amd.engine import default_run
    m, _, prob, pb = setup
    h = eng.knot_steps(m, prob, default_run(h=H0, seed=3, adaptive_h=1))
    ref = _rule(m, pb)
    assert np.array_equal(h, ref), np.max(np.abs(h - ref))
    mc = ~m.on_boundary
    assert np.sum(h[mc] < H0) >= 8 and np.all(h[mc] > 0) and np.all(h[mc] <= H0)
    # without the rule every knot takes the run's h
    assert np.all(eng.knot_steps(m, prob, default_run(h=H0, seed=3))[mc] == H0)


@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
def test_adaptive_knot_bit_exact(eng, setup, plain, lut_lds):
    from pdds_amd.engine import default_run
    m, fac, prob, pb = setup
    run = default_run(h=H0, seed=5, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds, adaptive_h=1)
    h = eng.knot_steps(m, prob, run)
    knot = int(np.argmin(np.where(m.on_boundary, np.inf, h)))
    assert h[knot] < 0.75 * H0
    n = 2000
    st = m.stencil(knot)
    print("adaptive h knot %d h_k %.3e plain=%d lut_lds=%d" % (knot, h[knot], plain, lut_lds), flush=True)
    got = eng.solve_knot(m, prob, run, knot, n)
    paths = O.simulate(pb, O.run_params(h[knot], 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, [fac[min(max(int(k), 0), 3)] for k in st["kind"]],
                           st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())


def test_adaptive_assembly_rows_equal_single_knots(eng, setup):
    from pdds_amd.engine import default_run
    m, _, prob, _ = setup
    run = default_run(h=H0, seed=7, adaptive_h=1)
    h = eng.knot_steps(m, prob, run)
    a = eng.assemble(m, prob, run, 500)
    fixed = eng.assemble(m, prob, default_run(h=H0, seed=7), 500)
    assert a["counters"]["path_steps"] > fixed["counters"]["path_steps"]
    mc = np.nonzero(~m.on_boundary)[0]
    small = int(mc[np.argmin(h[mc])])
    full = int(mc[np.argmax(h[mc])])
    for q in (small, full):
        one = eng.solve_knot(m, prob, run, q, 500)
        lo, hi = m.row_ptr[q], m.row_ptr[q + 1]
        assert np.array_equal(a["G"][lo:hi], one["G"]) and a["B"][q] == one["B"], q
    # FP32 paths take the same per-knot steps
    f32 = eng.assemble(m, prob, default_run(h=H0, seed=7, adaptive_h=1, precision="fp32"), 500)
    assert abs(f32["counters"]["path_steps"] / a["counters"]["path_steps"] - 1.0) < 0.02
