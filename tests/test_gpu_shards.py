"""Shard-count invariance on one GPU, with logical shards.

A split by knots -- the form a multi-GPU run would take -- assembles contiguous slices of the knot
list in separate calls (the other knots at zero paths).  Philox streams are keyed by (knot, path)
and every reduction has a fixed order, so the slices' rows put together must equal one assembly
of all knots bit for bit, for 2, 3 and 5 shards, in the LDS-tiled LUT kernel (whose tile patches
change with the slice) and the persistent kernel of an analytic PDE.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _problem(eng, kind):
    from pdds_amd.engine import bvp
    from pdds_amd.workloads import varcoef_fields
    if kind == "sinsin":
        return bvp("sinsin")
    xs = np.linspace(-1, 1, 96)
    X, Y = np.meshgrid(xs, xs)
    return bvp("varcoef_lut", lut=eng.lut(varcoef_fields(X, Y), (-1, -1, 1, 1), mode=0))


@pytest.mark.parametrize("kind", ["varcoef_lut", "sinsin"], ids=["lut-tiled", "sinsin-persistent"])
def test_logical_shards_bit_identical(eng, kind):
    from pdds_amd.engine import Mesh, default_run
    from pdds_amd.workloads import relative_fac
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    m = Mesh(sw, ne, 4, 8, relative_fac(sw, ne, 4, 8, 1.0))
    prob = _problem(eng, kind)
    run = default_run(h=1e-3, seed=29, sub_batch=32, item_paths=640)
    n = 1200
    full = eng.assemble(m, prob, run, n)
    rp = m.row_ptr
    for shards in (2, 3, 5):
        print("%s: %d logical shards" % (kind, shards), flush=True)
        G, Gv = np.full_like(full["G"], np.nan), np.full_like(full["Gvar"], np.nan)
        B, BB = np.full(m.n_knots, np.nan), np.full(m.n_knots, np.nan)
        steps = 0
        cuts = np.linspace(0, m.n_knots, shards + 1).astype(int)
        for k0, k1 in zip(cuts[:-1], cuts[1:]):
            N = np.zeros(m.n_knots, dtype=np.int64)
            N[k0:k1] = n
            out = eng.assemble(m, prob, run, 0, paths_per_knot=N)
            # knots outside the slice come back as zero rows
            outside = np.ones(m.n_knots, bool)
            outside[k0:k1] = False
            outside &= ~m.on_boundary
            for k in np.flatnonzero(outside):
                assert not np.any(out["G"][rp[k]:rp[k + 1]]) and out["B"][k] == 0.0
            G[rp[k0]:rp[k1]] = out["G"][rp[k0]:rp[k1]]
            Gv[rp[k0]:rp[k1]] = out["Gvar"][rp[k0]:rp[k1]]
            B[k0:k1] = out["B"][k0:k1]
            BB[k0:k1] = out["BB"][k0:k1]
            steps += out["counters"]["path_steps"]
        assert np.array_equal(G, full["G"]) and np.array_equal(Gv, full["Gvar"]), shards
        assert np.array_equal(B, full["B"]) and np.array_equal(BB, full["BB"]), shards
        assert steps == full["counters"]["path_steps"], shards
