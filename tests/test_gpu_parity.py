"""GPU parity: the HIP engine against the CPU oracle on identical Philox streams.

FP64 path state is bit-exact with the oracle (same operation order, no contraction,
same reproducible elementary functions), so per-path records, exit sides, step
counts and every G/B sum are compared with ==.  FP32 path state is compared
statistically with the FP64 oracle on the same random stream.
"""
import numpy as np
import pytest

from conftest import FIX_FAC, FIX_H, FIX_N, FIX_NNODE, FIX_NSUB, FIX_NE, FIX_SW, zscores
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _product_mesh(sw=FIX_SW, ne=FIX_NE, nsub=FIX_NSUB, nnode=FIX_NNODE, fac=FIX_FAC):
    from pdds_amd.engine import Mesh
    return Mesh(sw, ne, nsub, nnode, fac)


def _cshape(st, fac):
    return [fac[min(max(int(k), 0), 3)] for k in st["kind"]]


CASES = [
    # (pde, k2, domain, nsub, nnode, knot, h)
    ("monegros", 0.0, (FIX_SW, FIX_NE), 4, 8, 40, 0.005),
    ("sinsin", 0.0, ((-1.0, -1.0), (1.0, 1.0)), 2, 16, 5, 1e-3),
    ("varcoef", 0.0, ((-1.0, -1.0), (1.0, 1.0)), 4, 8, 30, 1e-3),
    ("helmholtz", 9.0, ((-1.0, -1.0), (1.0, 1.0)), 4, 8, 12, 1e-3),
]
PDE_ID = {"monegros": O.PDE_MONEGROS, "sinsin": O.PDE_SINSIN, "varcoef": O.PDE_VARCOEF,
          "helmholtz": O.PDE_HELMHOLTZ}


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_paths_bit_exact(eng, case):
    from pdds_amd.engine import bvp, default_run
    pde, k2, dom, nsub, nnode, knot, h = case
    m = _product_mesh(dom[0], dom[1], nsub, nnode, FIX_FAC)
    st = m.stencil(knot)
    n = 3000
    run = default_run(h=h, seed=7)
    got = eng.paths(bvp(pde, k2=k2), run, m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.simulate(O.problem(PDE_ID[pde], k2=k2), O.run_params(h, 7, threads=8), m.x0[knot], m.y0[knot],
                     st["rect"], knot, 0, n)
    for k in ("steps", "xt", "yt", "Y", "Z", "xi"):
        assert np.array_equal(got[k], ref[k]), (pde, k, np.flatnonzero(got[k] != ref[k])[:5])
    assert got["counters"]["path_steps"] == int(ref["steps"].sum())


def _lut_pair(eng, n=96):
    """The same 5-field coefficient table on the device and in the oracle."""
    from pdds_amd.workloads import varcoef_fields
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    samples = varcoef_fields(X, Y)
    lut = eng.lut(samples, (-1, -1, 1, 1), mode=0)
    d = 2.0 / (n - 1)
    olut = {"coef": O.lut_prefilter(samples), "NX": n + 2, "NY": n + 2, "nf": 5, "ox": -1 - d, "oy": -1 - d,
            "sx": 1 / d, "sy": 1 / d}
    return lut, olut


@pytest.mark.parametrize("precision", ["fp64"])
def test_fp64_lut_paths_bit_exact(eng, precision):
    from pdds_amd.engine import bvp, default_run
    lut, olut = _lut_pair(eng)
    m = _product_mesh((-1.0, -1.0), (1.0, 1.0), 4, 8, FIX_FAC)
    k = 30
    st = m.stencil(k)
    got = eng.paths(bvp("varcoef_lut", lut=lut), default_run(h=1e-3, seed=9), m.x0[k], m.y0[k], st["rect"], k, 0, 2000)
    ref = O.simulate(O.problem(O.PDE_VARCOEF_LUT, lut=olut), O.run_params(1e-3, 9, threads=8), m.x0[k], m.y0[k],
                     st["rect"], k, 0, 2000)
    for key in ("steps", "xt", "yt", "Y", "Z", "xi"):
        assert np.array_equal(got[key], ref[key]), key


def test_fp64_phase2_bit_exact(eng):
    from pdds_amd.engine import bvp, default_run
    m = _product_mesh()
    k = 40
    st = m.stencil(k)
    got = eng.paths(bvp("monegros"), default_run(h=0.005, seed=3), m.x0[k], m.y0[k], st["rect"], k, 100, 1500,
                    phase2=True)
    ref = O.simulate(O.problem(O.PDE_MONEGROS), O.run_params(0.005, 3, phase2=1, threads=8), m.x0[k], m.y0[k],
                     st["rect"], k, 100, 1500, phase2_out=True)
    for key in ("xt", "yt", "Y", "Z", "xi", "steps"):
        assert np.array_equal(got["phase2"][key], ref["phase2"][key]), key
    assert (got["phase2"]["steps"] >= 1).all()


@pytest.mark.parametrize("quirk", [1, 0])
@pytest.mark.parametrize("knot", [1, 40, 100])
def test_solve_knot_bit_exact(eng, knot, quirk):
    """One knot's G row / variance / B / BB: engine == oracle with the engine's summation order."""
    from pdds_amd.engine import bvp, default_run
    m = _product_mesh()
    st = m.stencil(knot)
    n = 2500
    run = default_run(h=FIX_H, seed=11, quirk_last_side=quirk, sub_batch=32, item_paths=640)
    got = eng.solve_knot(m, bvp("monegros"), run, knot, n)
    pb = O.problem(O.PDE_MONEGROS)
    paths = O.simulate(pb, O.run_params(FIX_H, 11, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, _cshape(st, FIX_FAC), st["ipsi"], 32, 640, quirk=quirk)
    assert np.array_equal(got["cols"], ref["cols"])
    assert np.array_equal(got["G"], ref["G"])
    assert np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["exits"] == [int(v) for v in ref["cnt"]]


@pytest.mark.parametrize("S,P", [(64, 1280), (48, 960), (8, 200)])
def test_solve_knot_sub_batch_sizes_bit_exact(eng, S, P):
    """Sub-batch sizes other than 32: with 48 or 64 paths a group of 8 sub-batches holds more
    records than the reduction has threads (the records are then loaded in a loop, not one per
    thread ahead of the group); with 8 the group is a quarter of the threads."""
    from pdds_amd.engine import bvp, default_run
    m = _product_mesh()
    knot, n = 40, 2500
    st = m.stencil(knot)
    print("sub-batch %d, item %d" % (S, P), flush=True)
    run = default_run(h=FIX_H, seed=11, sub_batch=S, item_paths=P)
    got = eng.solve_knot(m, bvp("monegros"), run, knot, n)
    pb = O.problem(O.PDE_MONEGROS)
    paths = O.simulate(pb, O.run_params(FIX_H, 11, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, _cshape(st, FIX_FAC), st["ipsi"], S, P, quirk=1)
    assert np.array_equal(got["cols"], ref["cols"])
    assert np.array_equal(got["G"], ref["G"])
    assert np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]


@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
def test_solve_knot_lut_bit_exact(eng, plain, lut_lds):
    """The assembly path kernels on the coefficient-LUT problem (the benchmark's kernels): the
    LDS-tiled kernel, the persistent kernel reading the LUT from global memory, and the plain
    launch: engine == oracle bit for bit."""
    from pdds_amd.engine import bvp, default_run
    from pdds_amd.workloads import relative_fac
    lut, olut = _lut_pair(eng)
    fac = relative_fac((-1.0, -1.0), (1.0, 1.0), 4, 8, 1.0)
    m = _product_mesh((-1.0, -1.0), (1.0, 1.0), 4, 8, fac)
    knot, n = 43, 3000                  # four non-empty stencil sides
    st = m.stencil(knot)
    run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds)
    got = eng.solve_knot(m, bvp("varcoef_lut", lut=lut), run, knot, n)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut)
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, _cshape(st, fac), st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"])
    assert np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_assemble_lut_tiled_equals_global(eng, precision):
    """A whole mesh on the coefficient-LUT problem: the LDS-tiled kernel (knots regrouped by
    tile, one queue per tile) gives exactly the G, G variances, B and BB of the kernel that reads
    the LUT from global memory, knot by knot."""
    from pdds_amd.engine import bvp, default_run
    from pdds_amd.workloads import relative_fac
    # 256^2 LUT on 4x4 subdomains: the interface-segment tiles (132 x 68 texels) fit in LDS,
    # the crossing-knot tiles (132 x 132) do not, so both patch kinds run in one launch
    lut, _ = _lut_pair(eng, n=256)
    m = _product_mesh((-1.0, -1.0), (1.0, 1.0), 4, 8, relative_fac((-1.0, -1.0), (1.0, 1.0), 4, 8, 1.0))
    outs = []
    for lds in (1, 0):
        run = default_run(h=2e-3, seed=21, precision=precision, lut_lds=lds)
        outs.append(eng.assemble(m, bvp("varcoef_lut", lut=lut), run, 1500))
    a, b = outs
    for k in ("G", "Gvar", "B", "BB"):
        assert np.array_equal(a[k], b[k]), k
    assert a["counters"]["path_steps"] == b["counters"]["path_steps"]
    assert a["counters"]["exits"] == b["counters"]["exits"]


def test_assemble_fixture_statistical_parity(eng, golden):
    """Whole fixture run (165 knots, N=5000, h=0.005) against the reference's recorded G/B at 3 SE."""
    from pdds_amd.engine import bvp, default_run
    m = _product_mesh()
    out = eng.assemble(m, bvp("monegros"), default_run(h=FIX_H, seed=42), FIX_N)
    z = []
    for k in range(m.n_knots):
        cols = m.row_cols(k)
        assert list(cols) == golden["rows"].get(k, [])
        sl = slice(m.row_ptr[k], m.row_ptr[k + 1])
        ref = [golden["G"][(k, int(c))] for c in cols]
        refv = [golden["Gvar"][(k, int(c))] for c in cols]
        z.extend(zscores(out["G"][sl], out["Gvar"][sl], FIX_N, ref, refv, FIX_N))
    z = np.array(z)
    assert len(z) == 6960
    assert abs(z.mean()) < 0.1 and 0.85 < z.std() < 1.2, (z.mean(), z.std())
    assert (np.abs(z) > 4.5).sum() == 0
    interior = ~m.on_boundary
    varB = out["BB"] - out["B"] ** 2
    zb = zscores(out["B"][interior], varB[interior], FIX_N,
                 [golden["B"][k] for k in np.flatnonzero(interior)],
                 [-golden["Bvar"][k] for k in np.flatnonzero(interior)], FIX_N)
    assert abs(zb.mean()) < 0.3 and zb.std() < 1.3 and np.abs(zb).max() < 4.5
    for k in np.flatnonzero(m.on_boundary):
        assert abs(out["B"][k] - golden["B"][k]) < 1e-5
    c = out["counters"]
    assert c["path_steps"] > 0 and c["capped"] == 0 and c["exits"][5] == 0


def test_fp32_against_fp64_oracle_same_streams(eng):
    """FP32 path state vs the FP64-stepped oracle on the same Philox draws (24-bit uniforms)."""
    from pdds_amd.engine import bvp, default_run
    m = _product_mesh((-1.0, -1.0), (1.0, 1.0), 2, 16, FIX_FAC)
    k = 5
    st = m.stencil(k)
    n = 20000
    got = eng.paths(bvp("sinsin"), default_run(h=1e-3, seed=42, precision="fp32"), m.x0[k], m.y0[k],
                    st["rect"], k, 0, n)
    ref = O.simulate(O.problem(O.PDE_SINSIN), O.run_params(1e-3, 42, normals=1, threads=8), m.x0[k], m.y0[k],
                     st["rect"], k, 0, n)
    same_len = np.mean(got["steps"] == ref["steps"])
    assert same_len > 0.97, same_len
    sel = got["steps"] == ref["steps"]
    assert np.abs(got["xt"][sel] - ref["xt"][sel]).max() < 5e-4
    assert np.abs(got["Z"][sel] - ref["Z"][sel]).max() < 1e-4
    assert np.abs(got["xi"][sel] - ref["xi"][sel]).max() < 1e-4
    # the estimators agree far inside their standard errors
    assert abs(got["Z"].mean() - ref["Z"].mean()) < 0.1 * got["Z"].std() / np.sqrt(n)


def test_point_eval_matches_oracle(eng):
    from pdds_amd.engine import bvp, default_run
    glob = np.array([-1.0, -1.0, 1.0, 1.0])
    x = np.array([0.1, -0.5, 0.7, 1.0])
    y = np.array([0.2, 0.3, -0.6, 0.0])
    out = eng.point_eval(bvp("sinsin"), default_run(h=1e-3, seed=5, item_paths=512), glob, x, y, 4096)
    pb = O.problem(O.PDE_SINSIN)
    for i in range(3):
        paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), x[i], y[i], glob, i, 0, 4096)
        st = {"rect": glob, "msz": np.zeros(4, dtype=np.int32), "idx": np.zeros(0, dtype=np.int32),
              "px": np.zeros(0), "py": np.zeros(0)}
        ref = O.reduce_batched(pb, paths, st, glob, [1.0] * 4, [np.zeros((0, 0))] * 4, 32, 512)
        assert out["u"][i] == ref["B"]
    assert out["u"][3] == bvp("sinsin").u(1.0, 0.0)
    u_exact = bvp("sinsin").u(x[:3], y[:3])
    assert np.all(np.abs(out["u"][:3] - u_exact) < 5 * np.sqrt(out["u_var"][:3] / 4096) + 0.02)


def test_device_bicgstab_matches_dense(eng):
    rng = np.random.default_rng(3)
    m = _product_mesh()
    vals = -np.abs(rng.normal(0, 0.02, m.nnz))
    B = rng.normal(3, 0.5, m.n_knots)
    u, it, rr = eng.solve(m.n_knots, m.row_ptr, m.cols, vals, B, tol=1e-13)
    A = np.eye(m.n_knots)
    for k in range(m.n_knots):
        A[k, m.row_cols(k)] += vals[m.row_ptr[k]:m.row_ptr[k + 1]]
    ud = np.linalg.solve(A, B)
    assert rr < 1e-12 and it > 0
    assert np.abs(u - ud).max() < 1e-10 * np.abs(ud).max()
    uo, ito, rro = O.bicgstab(np.asarray(m.row_ptr), m.cols, vals, B, tol=1e-13)
    assert np.abs(uo - ud).max() < 1e-10 * np.abs(ud).max()
    # same iteration as the oracle up to the order of the dot-product sums
    assert abs(it - ito) <= 2


def test_device_bicgstab_zero_rhs(eng):
    m = _product_mesh()
    vals = np.full(m.nnz, -0.01)
    u, it, rr = eng.solve(m.n_knots, m.row_ptr, m.cols, vals, np.zeros(m.n_knots), tol=1e-12)
    assert it == 0 and rr == 0.0 and not np.any(u)


def test_device_bicgstab_large_mesh(eng):
    """The 8x8 x 32-knot mesh (3437 knots, ~8.7e5 entries): G with the sign and row sums of an
    interface system (G <= 0, row sums in (-1, 0)), against a dense solve, twice in a row (the
    second call reuses the device buffers)."""
    from pdds_amd import workloads as W
    from pdds_amd.engine import Mesh
    sw, ne, ns, nn, _, _, _, _, shape = W.CONFIGS["cfg2_varcoef_lut"]
    m = Mesh(sw, ne, ns, nn, W.relative_fac(sw, ne, ns, nn, shape))
    rng = np.random.default_rng(11)
    vals = -rng.uniform(0.0, 1.0, m.nnz)
    for k in range(m.n_knots):
        a, b = m.row_ptr[k], m.row_ptr[k + 1]
        if b > a:
            vals[a:b] *= rng.uniform(0.2, 0.9) / vals[a:b].sum() * -1.0
    A = np.eye(m.n_knots)
    for k in range(m.n_knots):
        A[k, m.row_cols(k)] += vals[m.row_ptr[k]:m.row_ptr[k + 1]]
    for seed in (1, 2):
        B = np.random.default_rng(seed).normal(3, 0.5, m.n_knots)
        u, it, rr = eng.solve(m.n_knots, m.row_ptr, m.cols, vals, B, tol=1e-12)
        ud = np.linalg.solve(A, B)
        assert rr < 1e-12 and it > 0
        assert np.abs(u - ud).max() < 1e-9 * np.abs(ud).max()


def test_lut_device_eval_matches_oracle(eng):
    nx, ny, nf = 64, 48, 5
    xs = np.linspace(-1, 1, nx)
    ys = np.linspace(-1, 1, ny)
    X, Y = np.meshgrid(xs, ys)
    samples = np.stack([1 + 0.3 * np.sin(X), 1 + 0.3 * np.cos(Y), 0.2 * Y, -0.2 * X, -(1 + X * X)], axis=-1)
    lut = eng.lut(samples, (-1, -1, 1, 1), mode=0)
    coef = O.lut_prefilter(samples)
    dx, dy = 2.0 / (nx - 1), 2.0 / (ny - 1)
    olut = {"coef": coef, "NX": nx + 2, "NY": ny + 2, "nf": nf, "ox": -1 - dx, "oy": -1 - dy,
            "sx": 1 / dx, "sy": 1 / dy}
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut)
    rng = np.random.default_rng(0)
    px = rng.uniform(-1.2, 1.2, 4000)
    py = rng.uniform(-1.2, 1.2, 4000)
    for f in range(nf):
        dev = lut.eval(f, px, py)
        ref = np.array([O.lib().orc_lut_eval(pb, f, a, b) for a, b in zip(px, py)])
        assert np.array_equal(dev, ref), f
    # node interpolation of the prefiltered spline
    dev = lut.eval(0, X.ravel(), Y.ravel())
    assert np.abs(dev - samples[..., 0].ravel()).max() < 2e-6
