"""Batching form: knot-equation statistics, variance-targeted per-knot path counts, autotuner.

- pdds_assemble_eq's equation mean and variance equal the oracle's, summed in the same order,
  bit for bit (FP64 paths are bit-exact, so every sum is).
- pdds_solve_adaptive meets its standard-error target with uneven counts, every round draws
  fresh paths, and its solution is exactly the plain assembly-and-solve with the counts and the
  Philox stream it returns.
- pdds_autotune only picks launch shapes that leave every result bit unchanged.
"""
import numpy as np
import pytest

from conftest import FIX_FAC, FIX_H, FIX_NE, FIX_NNODE, FIX_NSUB, FIX_SW
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _cshape(st, fac):
    return [fac[min(max(int(k), 0), 3)] for k in st["kind"]]


@pytest.mark.parametrize("quirk", [1, 0])
def test_assemble_eq_bit_exact_with_oracle(eng, quirk):
    from pdds_amd.engine import Mesh, bvp, default_run
    m = Mesh(FIX_SW, FIX_NE, FIX_NSUB, FIX_NNODE, FIX_FAC)
    prob = bvp("monegros")
    u_ref = prob.u(m.px, m.py) + 0.01 * np.cos(np.arange(m.n_knots))
    n = 2500
    run = default_run(h=FIX_H, seed=13, quirk_last_side=quirk, sub_batch=32, item_paths=640)
    got = eng.assemble_eq(m, prob, run, n, u_ref)
    plain = eng.assemble(m, prob, run, n)
    # the equation statistics change nothing else
    for k in ("G", "Gvar", "B", "BB"):
        assert np.array_equal(got[k], plain[k]), k
    pb = O.problem(O.PDE_MONEGROS)
    for k in (1, 40, 100):
        st = m.stencil(k)
        paths = O.simulate(pb, O.run_params(FIX_H, 13, threads=8), m.x0[k], m.y0[k], st["rect"], k, 0, n)
        ref = O.reduce_batched_eq(pb, paths, st, m.glob, _cshape(st, FIX_FAC), st["ipsi"], 32, 640, u_ref,
                                  quirk=quirk)
        assert got["eq_mean"][k] == ref["eq_mean"], k
        assert got["eq_var"][k] == ref["eq_var"], k
        assert got["B"][k] == ref["B"]
    for k in np.flatnonzero(m.on_boundary):
        assert got["eq_var"][k] == 0.0 and got["eq_mean"][k] == got["B"][k]


@pytest.mark.parametrize("variance,target", [(0, 2e-3), (1, 5e-4)], ids=["score", "equation"])
def test_adaptive_counts_meet_the_target(eng, variance, target):
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    m = Mesh(sw, ne, 2, 16, relative_fac(sw, ne, 2, 16, 1.0))
    prob = bvp("sinsin")
    run = default_run(h=1e-3, seed=42)
    out = eng.solve_adaptive(m, prob, run, se_target=target, n_min=1000, n_max=1000000, max_rounds=4,
                             variance=variance)
    rep, N, se = out["report"], out["paths_per_knot"], out["se"]
    mc = ~m.on_boundary
    print(rep, np.sort(N[mc])[[0, len(N[mc]) // 2, -1]], flush=True)
    # the final standard errors are measured on the final round's own paths: within the target up
    # to the sampling error of a batch-means variance (relative sd about sqrt(1/(2 batches)) for the
    # variance that set the count and again for the final one)
    nb = np.ceil(N[mc] / 32)
    assert rep["knots_capped"] == 0 and rep["knots_over"] == int(np.sum(se[mc] > target)), rep
    assert np.all(se[mc] <= target * (1 + 5 * np.sqrt(1 / nb))), np.max(se[mc] / target)
    assert rep["knots_over"] <= 0.1 * mc.sum() and np.isclose(rep["max_se"], se[mc].max())
    assert np.all(N[mc] >= 1000) and np.all(N[~mc] == 0)
    assert N[mc].max() > 1.5 * N[mc].min(), (N[mc].min(), N[mc].max())   # the counts follow the variances
    assert rep["paths"] == int(N.sum()) and rep["path_steps"] > 0
    # every round draws fresh paths (a stream of its own), and the result is one plain assembly with
    # the final counts on the final round's stream, solved on the device
    assert rep["stream"] == (run.stream + rep["rounds"] * 0x9E3779B9) % 2 ** 32 != run.stream
    run_final = default_run(h=1e-3, seed=42, stream=rep["stream"])
    a = eng.assemble(m, prob, run_final, 0, paths_per_knot=np.where(mc, N, 1))
    u, it, rr = eng.solve(m.n_knots, m.row_ptr, m.cols, a["G"], a["B"], tol=1e-10)
    assert np.array_equal(u, out["u"]) and it == rep["iters"]
    err = u - prob.u(m.px, m.py)
    assert np.sqrt(np.mean(err ** 2)) < 5e-3


def test_autotune_leaves_results_unchanged(eng):
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac, varcoef_fields
    xs = np.linspace(-1, 1, 96)
    X, Y = np.meshgrid(xs, xs)
    lut = eng.lut(varcoef_fields(X, Y), (-1, -1, 1, 1), mode=0)
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    m = Mesh(sw, ne, 4, 8, relative_fac(sw, ne, 4, 8, 1.0))
    prob = bvp("varcoef_lut", lut=lut)
    base = default_run(h=1e-3, seed=3)
    best, rep = eng.autotune(m, prob, base, sample_paths=2048, max_knots=64, reps=1)
    assert rep["candidates"] >= 10 and rep["best_rate"] >= rep["base_rate"] > 0
    assert best.h == base.h and best.seed == base.seed and best.precision == base.precision
    a = eng.assemble(m, prob, base, 3000)
    b = eng.assemble(m, prob, best, 3000)
    for k in ("G", "Gvar", "B", "BB"):
        assert np.array_equal(a[k], b[k]), k
    assert a["counters"]["path_steps"] == b["counters"]["path_steps"]
