"""Ragged per-knot path counts and the step cap, bit-exact with the oracle.

One assembly carries Monte Carlo knots with 1, 2, 31, 32, 33, 639, 640, 641 and 1500 paths
(sub-batches of 32, work items of 640 paths), so single-path knots, partial sub-batches and item
boundaries meet in one launch; every such knot's G row, variances, B and BB equal the oracle's
engine-order reduction of the same Philox paths, in the persistent and in the plain kernel.  With
a step cap of 40 steps many paths stop at the cap; every sum still equals the oracle's, which
records the same projection for a capped path.
"""
import numpy as np
import pytest

from conftest import FIX_FAC, FIX_H, FIX_NE, FIX_NNODE, FIX_NSUB, FIX_SW
from oracle import oracle as O

pytestmark = pytest.mark.gpu

COUNTS = [1, 2, 31, 32, 33, 639, 640, 641, 1500]


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _cshape(st, fac):
    return [fac[min(max(int(k), 0), 3)] for k in st["kind"]]


def _oracle_row(m, k, n, seed, max_steps=0):
    pb = O.problem(O.PDE_MONEGROS)
    st = m.stencil(k)
    paths = O.simulate(pb, O.run_params(FIX_H, seed, max_steps=max_steps, threads=8), m.x0[k], m.y0[k],
                       st["rect"], k, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, _cshape(st, FIX_FAC), st["ipsi"], 32, 640, quirk=1)
    return ref, paths


@pytest.mark.parametrize("plain", [0, 1], ids=["persistent", "plain"])
def test_ragged_counts_bit_exact(eng, plain):
    from pdds_amd.engine import Mesh, bvp, default_run
    m = Mesh(FIX_SW, FIX_NE, FIX_NSUB, FIX_NNODE, FIX_FAC)
    mc = np.flatnonzero(~m.on_boundary)
    picks = mc[np.linspace(0, len(mc) - 1, len(COUNTS)).astype(int)]
    N = np.full(m.n_knots, 3, dtype=np.int64)
    N[picks] = COUNTS
    run = default_run(h=FIX_H, seed=17, sub_batch=32, item_paths=640, plain=plain)
    print("ragged counts, plain=%d, knots %s" % (plain, list(picks)), flush=True)
    out = eng.assemble(m, bvp("monegros"), run, 0, paths_per_knot=N)
    for k, c in zip(picks, COUNTS):
        ref, _ = _oracle_row(m, int(k), c, 17)
        a, b = m.row_ptr[k], m.row_ptr[k + 1]
        assert np.array_equal(m.cols[a:b], ref["cols"]), (k, c)
        assert np.array_equal(out["G"][a:b], ref["G"]), (k, c)
        assert np.array_equal(out["Gvar"][a:b], ref["Gvar"], equal_nan=True), (k, c)
        assert out["B"][k] == ref["B"] and out["BB"][k] == ref["BB"], (k, c)
    assert out["counters"]["paths"] == int(N[mc].sum())
    # the call's step count is the oracle's over every knot's paths: with 3 paths on most knots
    # a wave's lanes span many knots, whose step counters then take one update per lane
    pb = O.problem(O.PDE_MONEGROS)
    steps = 0
    for k in mc:
        st = m.stencil(int(k))
        p = O.simulate(pb, O.run_params(FIX_H, 17, threads=8), m.x0[k], m.y0[k], st["rect"], int(k), 0, int(N[k]))
        steps += int(p["steps"].sum())
    assert out["counters"]["path_steps"] == steps


def test_step_cap_bit_exact(eng):
    from pdds_amd.engine import Mesh, bvp, default_run
    m = Mesh(FIX_SW, FIX_NE, FIX_NSUB, FIX_NNODE, FIX_FAC)
    knot, n, cap = 40, 2000, 40
    run = default_run(h=FIX_H, seed=23, sub_batch=32, item_paths=640, max_steps=cap)
    print("step cap %d, knot %d, %d paths" % (cap, knot, n), flush=True)
    got = eng.solve_knot(m, bvp("monegros"), run, knot, n)
    ref, paths = _oracle_row(m, knot, n, 23, max_steps=cap)
    assert np.array_equal(got["G"], ref["G"])
    assert np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    at_cap = int(np.sum(paths["steps"] == cap))
    assert 0 < got["counters"]["capped"] <= at_cap and int(paths["steps"].max()) == cap
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())
