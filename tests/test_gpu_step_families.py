"""The step families of run.step on the GPU: plain (1, no control variate) and VR (2, the drift
change by mu = -sigma^T grad u / u) beside the default control-variate step.

On the coefficient-LUT problem, the LDS-tiled, persistent and plain kernels give a knot's G row,
variances, B and BB bit-exact with the oracle's step of the same family.  With the plain step G
depends only on the exit positions and Y, so it equals the control-variate run's G while B (the
score without xi) differs; pdds_paths returns xi = 0 with the control-variate run's X, Y, Z and
step counts.  The VR paths move differently, carry xi = 0, and their estimate of u at a point has
a small fraction of the plain step's variance.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _lut_pair(eng, n=96):
    from pdds_amd.workloads import varcoef_fields
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    samples = varcoef_fields(X, Y)
    lut = eng.lut(samples, (-1, -1, 1, 1), mode=0)
    d = 2.0 / (n - 1)
    olut = {"coef": O.lut_prefilter(samples), "NX": n + 2, "NY": n + 2, "nf": 5, "ox": -1 - d, "oy": -1 - d,
            "sx": 1 / d, "sy": 1 / d}
    return lut, olut


@pytest.mark.parametrize("step", [1, 2], ids=["plain-step", "vr-step"])
@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
def test_step_family_bit_exact(eng, plain, lut_lds, step):
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac
    lut, olut = _lut_pair(eng)
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    fac = relative_fac(sw, ne, 4, 8, 1.0)
    m = Mesh(sw, ne, 4, 8, fac)
    knot, n = 43, 3000
    st = m.stencil(knot)
    prob = bvp("varcoef_lut", lut=lut)
    print("step family %d, plain=%d lut_lds=%d, knot %d" % (step, plain, lut_lds, knot), flush=True)
    run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds, step=step)
    got = eng.solve_knot(m, prob, run, knot, n)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut, step=step)
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    assert np.all(paths["xi"] == 0.0)
    ref = O.reduce_batched(pb, paths, st, m.glob, [fac[min(max(int(k), 0), 3)] for k in st["kind"]],
                           st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())
    cv = eng.solve_knot(m, prob, default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain,
                                             lut_lds=lut_lds), knot, n)
    if step == 1:
        assert np.array_equal(got["G"], cv["G"]) and got["B"] != cv["B"]
    else:
        assert not np.array_equal(got["G"], cv["G"])


def test_plain_step_paths_have_no_xi(eng):
    from pdds_amd.engine import bvp, default_run
    prob = bvp("sinsin")
    rect = (-1.0, -1.0, 1.0, 1.0)
    a = eng.paths(prob, default_run(h=1e-3, seed=9, step=1), 0.3, -0.2, rect, 7, 0, 4096)
    b = eng.paths(prob, default_run(h=1e-3, seed=9), 0.3, -0.2, rect, 7, 0, 4096)
    assert np.all(a["xi"] == 0.0) and np.any(b["xi"] != 0.0)
    for k in ("xt", "yt", "Y", "Z", "steps"):
        assert np.array_equal(a[k], b[k]), k


def test_vr_point_estimate(eng):
    """VR at an interior point of the LUT problem (u = e^x cos y > 0): unbiased within the step's
    bias, xi = 0, and a small fraction of the plain step's variance."""
    from pdds_amd.engine import bvp, default_run
    lut, _ = _lut_pair(eng)
    prob = bvp("varcoef_lut", lut=lut)
    glob = (-1.0, -1.0, 1.0, 1.0)
    x, y = np.array([0.3]), np.array([-0.2])
    est = {}
    for step in (1, 2):
        est[step] = eng.point_eval(prob, default_run(h=1e-3, seed=21, step=step), glob, x, y, 200000)
    u = np.exp(0.3) * np.cos(-0.2)
    u_vr, var_vr = est[2]["u"][0], est[2]["u_var"][0]
    assert abs(u_vr - u) < 4 * np.sqrt(var_vr / 200000) + 12e-3, (u_vr, u)
    assert var_vr < 0.05 * est[1]["u_var"][0], (var_vr, est[1]["u_var"][0])
    a = eng.paths(prob, default_run(h=1e-3, seed=9, step=2), 0.3, -0.2, glob, 7, 0, 1024)
    assert np.all(a["xi"] == 0.0)
