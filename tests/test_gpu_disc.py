"""Disc meshes on the GPU: annular-sector stopping domains in every path kernel.

FP64 knots of a disc mesh are bit-exact with the oracle on the same Philox stream (paths, exit
sides, every G / B sum) in the LDS-tiled, persistent and plain kernels; the variance-targeted
solve on a disc meets its target and approaches the manufactured solution.
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


@pytest.fixture(scope="module")
def disc():
    from pdds_amd.engine import Mesh
    return Mesh.disc(1.0, 6, 24, 11, 8, 64, 1.0)


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


CASES = [("varcoef_lut", 0, 1), ("varcoef_lut", 0, 0), ("varcoef_lut", 1, 0), ("sinsin", 0, 0), ("sinsin", 1, 0)]


@pytest.mark.parametrize("pde,plain,lut_lds", CASES,
                         ids=["lut-tiled", "lut-persistent", "lut-plain", "sinsin-persistent", "sinsin-plain"])
def test_disc_solve_knot_bit_exact(eng, disc, pde, plain, lut_lds):
    from pdds_amd.engine import bvp, default_run
    m = disc
    if pde == "varcoef_lut":
        lut, olut = _lut_pair(eng)
        prob, pb, h = bvp(pde, lut=lut), O.problem(O.PDE_VARCOEF_LUT, lut=olut), 2e-4
    else:
        prob, pb, h = bvp(pde), O.problem(O.PDE_SINSIN), 1e-4
    n = 3000
    run = default_run(h=h, seed=19, plain=plain, lut_lds=lut_lds, sub_batch=32, item_paths=640)
    crossing = m.n_knots - 24 - 60
    for k in (5, 2000, 7000, crossing):
        st = m.stencil(k)
        got = eng.solve_knot(m, prob, run, k, n)
        paths = O.simulate(pb, O.run_params(h, 19, threads=8), m.x0[k], m.y0[k], st["domain"], k, 0, n, geom=1)
        # disc meshes sum a corner entry's sides whatever quirk_last_side says
        ref = O.reduce_batched_eq(pb, paths, st, m.glob, st["cshape"], st["ipsi"], 32, 640, None, quirk=0, geom=1)
        assert np.array_equal(got["cols"], ref["cols"]), k
        assert np.array_equal(got["G"], ref["G"]), k
        assert np.array_equal(got["Gvar"], ref["Gvar"]), k
        assert got["B"] == ref["B"] and got["BB"] == ref["BB"], k
        assert got["counters"]["exits"] == [int(v) for v in ref["cnt"]], k
        assert got["counters"]["exits"][5] == 0      # every exit classified


def test_disc_adaptive_solve(eng):
    """Variance-targeted counts on a disc meet the target up to the sampling error of the measured
    variances; each knot equation, weighed with the
    exact u, reproduces u at the knot within its error; the solved interface values are close to
    u."""
    from pdds_amd.engine import Mesh, bvp, default_run
    m = Mesh.disc(1.0, 4, 12, 11, 8, 16, 1.0)
    prob = bvp("sinsin")
    run = default_run(h=2e-4, seed=5)
    target = 4e-3
    out = eng.solve_adaptive(m, prob, run, se_target=target, n_min=1000, n_max=1000000, max_rounds=4)
    rep, N, se = out["report"], out["paths_per_knot"], out["se"]
    mc = ~m.on_boundary
    # standard errors (of each knot's score) measured on the final round's own paths: the target up
    # to the sampling error of the batch-means variances
    assert rep["knots_capped"] == 0 and rep["knots_over"] <= 0.1 * mc.sum(), rep
    assert np.all(se[mc] <= target * (1 + 5 * np.sqrt(1 / np.ceil(N[mc] / 32)))), np.max(se[mc]) / target
    assert np.all(N[mc] >= 1000) and rep["paths"] == int(N.sum())
    uex = prob.u(m.px, m.py)
    eq = eng.assemble_eq(m, prob, run, 0, uex, paths_per_knot=np.where(mc, N, 1))
    dev = np.abs(eq["eq_mean"] - uex)[mc]
    se = np.sqrt(eq["eq_var"][mc] / N[mc])
    assert np.mean(dev < 5 * se + 5e-3) > 0.97, np.sort(dev)[-5:]
    assert np.sqrt(np.mean((eq["eq_mean"][mc] - uex[mc]) ** 2)) < 3e-3
    err = out["u"] - uex
    rms = np.sqrt(np.mean(err ** 2))
    print(rep, rms, flush=True)
    assert rms < 2e-2, rms
