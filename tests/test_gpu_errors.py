"""Error behaviour of the engine's entry points on the GPU.

Each bad argument is turned into a status code and a message by the host code before any launch,
and the engine stays usable afterwards: a valid call following the failed ones returns the same
result as on a fresh engine.
"""
import numpy as np
import pytest

from pdds_amd import _capi as A

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def mesh():
    from pdds_amd.engine import Mesh
    from pdds_amd.workloads import relative_fac
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    return Mesh(sw, ne, 2, 8, relative_fac(sw, ne, 2, 8, 1.0))


def _status(fn):
    from pdds_amd.engine import PddsError
    with pytest.raises(PddsError) as ex:
        fn()
    return ex.value.status, str(ex.value)


@pytest.mark.parametrize("field,value,msg", [
    ("step", 4, "step family"),
    ("step", -1, "step family"),
    ("max_steps", 1 << 27, "max_steps"),
    ("precision", 2, "precision"),
    ("wg_size", 96, "wg_size"),
    ("paths_per_lane", 65, "paths_per_lane"),
    ("h", 0.0, "h must be positive"),
    ("item_paths", 100, "multiple of sub_batch"),
])
def test_bad_run_parameters(eng, mesh, field, value, msg):
    from pdds_amd.engine import bvp, default_run
    run = default_run(h=1e-3, seed=1, sub_batch=32)
    setattr(run, field, value)
    print("bad run parameter %s = %r" % (field, value), flush=True)
    st, text = _status(lambda: eng.assemble(mesh, bvp("sinsin"), run, 100))
    assert st == A.ERR_INVALID and msg in text, text


def test_bad_counts_and_problems(eng, mesh):
    from pdds_amd.engine import bvp, default_run
    run = default_run(h=1e-3, seed=1)
    # uniform count below one
    st, text = _status(lambda: eng.assemble(mesh, bvp("sinsin"), run, 0))
    assert st == A.ERR_INVALID and "n_paths" in text
    # a negative per-knot count
    N = np.full(mesh.n_knots, 10, dtype=np.int64)
    N[~mesh.on_boundary] = -1
    st, text = _status(lambda: eng.assemble(mesh, bvp("sinsin"), run, 0, paths_per_knot=N))
    assert st == A.ERR_INVALID and "negative" in text
    # a knot index out of range
    st, _ = _status(lambda: eng.solve_knot(mesh, bvp("sinsin"), run, -1, 100))
    assert st == A.ERR_INVALID
    # an empty path range is fine, a range past 2^32 is not
    rect = (-1.0, -1.0, 1.0, 1.0)
    assert eng.paths(bvp("sinsin"), run, 0.1, 0.1, rect, 0, 0, 0)["counters"]["paths"] == 0
    st, _ = _status(lambda: eng.paths(bvp("sinsin"), run, 0.1, 0.1, rect, 0, (1 << 32) - 1, 2))
    assert st == A.ERR_INVALID
    # tables of the wrong shape: a tabulated problem needs 8 + 1 fields, VR+CV a 4-field table
    xs = np.linspace(-1, 1, 8)
    X, Y = np.meshgrid(xs, xs)
    l5 = eng.lut(np.stack([X] * 5, axis=-1), rect)
    l1 = eng.lut(X, rect)
    st, text = _status(lambda: eng.paths(bvp("table", lut=l5, ulut=l1), run, 0.1, 0.1, rect, 0, 0, 8))
    assert st == A.ERR_INVALID and "8-field" in text
    st, text = _status(lambda: eng.paths(bvp("sinsin"), default_run(h=1e-3, seed=1, step=3), 0.1, 0.1, rect, 0, 0, 8))
    assert st == A.ERR_INVALID and "vlut" in text
    st, text = _status(lambda: eng.paths(bvp("sinsin", vlut=l5), default_run(h=1e-3, seed=1, step=3), 0.1, 0.1, rect,
                                         0, 0, 8))
    assert st == A.ERR_INVALID and "vlut" in text


def test_engine_usable_after_errors(eng, mesh):
    from pdds_amd.engine import Engine, bvp, default_run
    run = default_run(h=1e-3, seed=4)
    a = eng.assemble(mesh, bvp("sinsin"), run, 500)
    fresh = Engine(0)
    try:
        b = fresh.assemble(mesh, bvp("sinsin"), run, 500)
    finally:
        fresh.close()
    for k in ("G", "Gvar", "B", "BB"):
        assert np.array_equal(a[k], b[k]), k
