"""The tabulated problem (every coefficient from LUTs, the reference's texture-served MCinCUDA
texMode 4, Solvers/MCinCUDA.cu:453-478) and the VR+CV step (Integrators/FKACIntegrator.hpp:100-112)
on the GPU.

FP64 paths of every step family (control variate, plain, VR, VR+CV) in the per-path, persistent
and plain kernels equal the oracle's paths, and a knot's G row, variances, B and BB equal the
oracle's batched reduction, bit for bit.  A boundary knot's B is u from the host copy of the u
table, equal to the oracle's lookup.  FP32 VR+CV paths give an estimate of u within the step's
bias and a small fraction of the plain step's variance.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

BOX = (-1.0, -1.0, 1.0, 1.0)


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def tables(eng):
    from pdds_amd import workloads as W
    n = 64
    prob = W.table_problem(eng, n, BOX, theta=0.5)
    xs = np.linspace(-1.0, 1.0, n)
    X, Y = np.meshgrid(xs, xs)
    olut = O.lut_dict(W.varcoef_table_fields(X, Y), BOX)
    oulut = O.lut_dict(W.varcoef_u(X, Y), BOX)
    ovlut = O.lut_dict(W.varcoef_vrcv_fields(X, Y, 0.5), BOX)
    return prob, olut, oulut, ovlut


def _oracle_problem(tables, step):
    _, olut, oulut, ovlut = tables
    return O.problem(O.PDE_TABLE, lut=olut, ulut=oulut, vlut=ovlut, step=step)


@pytest.mark.parametrize("step", [0, 1, 2, 3], ids=["cv", "plain-step", "vr", "vrcv"])
def test_table_paths_bit_exact(eng, tables, step):
    from pdds_amd.engine import default_run
    prob = tables[0]
    rect = (-0.5, -0.6, 0.6, 0.5)
    print("table paths, step family %d" % step, flush=True)
    got = eng.paths(prob, default_run(h=2e-3, seed=13, step=step), 0.05, -0.1, rect, 6, 0, 2048)
    ref = O.simulate(_oracle_problem(tables, step), O.run_params(2e-3, 13, threads=8), 0.05, -0.1,
                     np.array(rect), 6, 0, 2048)
    for k in ("xt", "yt", "Y", "Z", "xi", "steps"):
        assert np.array_equal(got[k], ref[k]), k
    if step in (1, 2):
        assert np.all(got["xi"] == 0.0)
    else:
        assert np.any(got["xi"] != 0.0)


@pytest.mark.parametrize("step", [0, 3], ids=["cv", "vrcv"])
@pytest.mark.parametrize("plain", [0, 1], ids=["persistent", "plain"])
def test_table_knot_bit_exact(eng, tables, step, plain):
    from pdds_amd.engine import Mesh, default_run
    from pdds_amd.workloads import relative_fac
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    fac = relative_fac(sw, ne, 4, 8, 1.0)
    m = Mesh(sw, ne, 4, 8, fac)
    knot, n = 43, 2000
    st = m.stencil(knot)
    print("table knot %d, step family %d, plain=%d" % (knot, step, plain), flush=True)
    run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain, step=step)
    got = eng.solve_knot(m, tables[0], run, knot, n)
    pb = _oracle_problem(tables, step)
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, [fac[min(max(int(k), 0), 3)] for k in st["kind"]],
                           st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())
    # a boundary knot: B = u from the host copy of the u table, the oracle's lookup
    onb = np.nonzero(m.on_boundary)[0][3]
    b = eng.solve_knot(m, tables[0], run, int(onb), n)
    assert b["B"] == O.pde_eval(pb, m.x0[onb], m.y0[onb])["g"]


def test_table_vrcv_fp32_estimate(eng, tables):
    from pdds_amd.engine import default_run
    prob = tables[0]
    rect = (-0.5, -0.6, 0.6, 0.5)
    u0 = np.exp(0.05) * np.cos(-0.1)
    est = {}
    for step in (1, 3):
        p = eng.paths(prob, default_run(h=1e-3, seed=17, step=step, precision="fp32"), 0.05, -0.1, rect, 3, 0, 16384)
        est[step] = p["Z"] + p["Y"] * np.exp(p["xt"]) * np.cos(p["yt"]) + p["xi"]
    s = est[3]
    assert abs(s.mean() - u0) < 4 * s.std() / np.sqrt(len(s)) + 6e-3, s.mean() - u0
    assert s.var() < 0.02 * est[1].var(), (s.var(), est[1].var())
