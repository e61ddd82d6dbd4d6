"""The plain Euler-Maruyama step (run.plain_step = 1, no control variate) on the GPU.

On the coefficient-LUT problem, the LDS-tiled, persistent and plain kernels give a knot's G row,
variances, B and BB bit-exact with the oracle's plain step. G depends only on the exit positions
and Y, so it equals the control-variate run's G. B (score without xi) differs from it, and
pdds_paths returns xi = 0 with the control-variate run's X, Y, Z and step counts.
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


@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
def test_plain_step_bit_exact(eng, plain, lut_lds):
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac
    lut, olut = _lut_pair(eng)
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    fac = relative_fac(sw, ne, 4, 8, 1.0)
    m = Mesh(sw, ne, 4, 8, fac)
    knot, n = 43, 3000
    st = m.stencil(knot)
    prob = bvp("varcoef_lut", lut=lut)
    print("plain step, plain=%d lut_lds=%d, knot %d" % (plain, lut_lds, knot), flush=True)
    run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds, plain_step=1)
    got = eng.solve_knot(m, prob, run, knot, n)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut, plain_step=1)
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    assert np.all(paths["xi"] == 0.0)
    ref = O.reduce_batched(pb, paths, st, m.glob, [fac[min(max(int(k), 0), 3)] for k in st["kind"]],
                           st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    cv = eng.solve_knot(m, prob, default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain,
                                             lut_lds=lut_lds), knot, n)
    assert np.array_equal(got["G"], cv["G"]) and got["B"] != cv["B"]


def test_plain_step_paths_have_no_xi(eng):
    from pdds_amd.engine import bvp, default_run
    prob = bvp("sinsin")
    rect = (-1.0, -1.0, 1.0, 1.0)
    a = eng.paths(prob, default_run(h=1e-3, seed=9, plain_step=1), 0.3, -0.2, rect, 7, 0, 4096)
    b = eng.paths(prob, default_run(h=1e-3, seed=9), 0.3, -0.2, rect, 7, 0, 4096)
    assert np.all(a["xi"] == 0.0) and np.any(b["xi"] != 0.0)
    for k in ("xt", "yt", "Y", "Z", "steps"):
        assert np.array_equal(a[k], b[k]), k
