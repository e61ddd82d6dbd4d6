"""pdds_point_eval_noalloc, called through the C ABI: the same results as pdds_point_eval, bit for
bit, on two calls in a row (the second smaller, so it reuses the scratch array the first one left
behind), for FP64 and FP32 paths and an LDS-tiled LUT problem; a call too large for the scratch
array is refused before any device work and leaves the engine usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DOMAIN = np.array([-1.0, -1.0, 1.0, 1.0])


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _noalloc(eng, problem, run, x, y, n_paths):
    from pdds_amd import _capi as A
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    npth = np.ascontiguousarray(np.broadcast_to(np.asarray(n_paths, dtype=np.int64), x.shape))
    u = np.full(len(x), np.nan)
    uv = np.full(len(x), np.nan)
    ctr = A.Counters()
    st = A.lib().pdds_point_eval_noalloc(eng.h, C.byref(problem._c()), C.byref(run), A.ptr(DOMAIN), len(x),
                                         A.ptr(x), A.ptr(y), A.ptr(npth, A._i64), A.ptr(u), A.ptr(uv), C.byref(ctr))
    return st, u, uv, ctr.as_dict()


def _problems(eng):
    from pdds_amd.engine import bvp
    from pdds_amd.workloads import varcoef_lut
    return {"sinsin": bvp("sinsin"), "helmholtz": bvp("helmholtz", k2=4.0),
            "varcoef_lut": bvp("varcoef_lut", lut=varcoef_lut(eng, 128))}


@pytest.mark.parametrize("pde,precision", [("sinsin", "fp64"), ("helmholtz", "fp32"), ("varcoef_lut", "fp64")])
def test_noalloc_equals_point_eval_twice(eng, pde, precision):
    from pdds_amd import _capi as A
    from pdds_amd.engine import default_run
    prob = _problems(eng)[pde]
    run = default_run(h=2e-3, seed=31, precision=precision, item_paths=512)
    rng = np.random.default_rng(17)
    for n_pts, n_paths in ((9, 60000), (4, 7000)):       # the second call is smaller
        x = rng.uniform(-0.95, 0.95, n_pts)
        y = rng.uniform(-0.95, 0.95, n_pts)
        x[0], y[0] = 1.0, 0.3                              # one point on the boundary
        print("noalloc %s %s: %d points x %d paths" % (pde, precision, n_pts, n_paths), flush=True)
        st, u, uv, ctr = _noalloc(eng, prob, run, x, y, n_paths)
        assert st == A.OK, A.lib().pdds_last_error(eng.h)
        ref = eng.point_eval(prob, run, DOMAIN, x, y, n_paths)
        assert np.array_equal(u, ref["u"]) and np.array_equal(uv, ref["u_var"])
        assert ctr["path_steps"] == ref["counters"]["path_steps"] > 0
        assert ctr["exits"] == ref["counters"]["exits"]


def test_noalloc_refuses_what_exceeds_the_scratch(eng):
    from pdds_amd import _capi as A
    from pdds_amd.engine import bvp, default_run
    run = default_run(h=2e-3, seed=31)
    st, u, uv, ctr = _noalloc(eng, bvp("sinsin"), run, [0.1, 0.2], [0.0, 0.1], 2_000_000)
    assert st == A.ERR_FULL
    assert b"no-allocation scratch" in A.lib().pdds_last_error(eng.h)
    # the engine and the scratch array remain usable
    st, u, uv, ctr = _noalloc(eng, bvp("sinsin"), run, [0.1], [0.0], 3000)
    assert st == A.OK and np.isfinite(u[0]) and ctr["path_steps"] > 0
