"""The benchmark and serving configurations against the oracle at their own sizes.

* The 8x8 x 32-knot mesh with 512^2 coefficient LUTs at h = 1e-4 (the benchmark workload):
  its largest segment stencil (314 entries, an LDS-tiled 0.5 x 0.25 patch) and a crossing
  knot (376 entries, whose 0.5 x 0.5 tile does not fit in LDS), in the tiled, persistent and
  plain kernels, bit-exact with the oracle; the FP32 engine on the same knots against the
  FP64-stepped oracle on the same Philox words; one whole assembly with its property checks.
* Serving rounds (FP64 paths): points of different streams and k^2 mixed with an interface
  solve in one launch (the per-lane stream and k^2 kernels), each point equal to the oracle on
  its own stream, and the solve equal to a direct assembly whose rows equal the oracle's.
* Raw-sample LUTs (mode 1, the reference's cubicTex2DSimple semantics) on the device.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG2_H, CFG2_SEED = 1e-4, 7
CFG2_KNOTS = {"segment-314": 281, "crossing-376": 311}


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def cfg2(eng):
    """The benchmark mesh and LUT problem, and the same LUT in the oracle."""
    from pdds_amd import workloads as W
    mesh, prob, n_paths, h, seed = W.build("cfg2_varcoef_lut", eng)
    assert (h, seed) == (CFG2_H, CFG2_SEED) and mesh.n_knots == 3437
    n = 512
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    d = 2.0 / (n - 1)
    olut = {"coef": O.lut_prefilter(W.varcoef_fields(X, Y)), "NX": n + 2, "NY": n + 2, "nf": 5, "ox": -1 - d,
            "oy": -1 - d, "sx": 1 / d, "sy": 1 / d}
    return mesh, prob, O.problem(O.PDE_VARCOEF_LUT, lut=olut)


def _cshape(mesh, k):
    st = mesh.stencil(k)
    return st, st["cshape"]


@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
@pytest.mark.parametrize("which", list(CFG2_KNOTS))
def test_cfg2_knot_bit_exact(eng, cfg2, which, plain, lut_lds):
    from pdds_amd.engine import default_run
    mesh, prob, opb = cfg2
    k, n = CFG2_KNOTS[which], 2000
    st, cs = _cshape(mesh, k)
    print("cfg2 knot %d (%d entries), plain=%d lut_lds=%d" % (k, int(st["msz"].sum()), plain, lut_lds), flush=True)
    run = default_run(h=CFG2_H, seed=CFG2_SEED, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds)
    got = eng.solve_knot(mesh, prob, run, k, n)
    paths = O.simulate(opb, O.run_params(CFG2_H, CFG2_SEED, threads=8), mesh.x0[k], mesh.y0[k], st["rect"], k, 0, n)
    ref = O.reduce_batched(opb, paths, st, mesh.glob, cs, st["ipsi"], 32, 640, quirk=1)
    assert len(ref["G"]) == mesh.row_nnz(k) and np.array_equal(got["cols"], ref["cols"])
    assert np.array_equal(got["G"], ref["G"])
    assert np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())
    assert got["counters"]["exits"] == [int(v) for v in ref["cnt"]]


def test_cfg2_assembly_rows_equal_oracle(eng, cfg2):
    """The benchmark's whole assembly -- 3409 Monte Carlo knots in one LDS-tiled launch, every
    patch and queue of it, FP64 paths -- at 2000 paths per knot: the G rows, their variances and
    the B entries of sampled knots (random ones, the largest segment stencil and a crossing knot)
    equal the oracle's for those knots bit for bit."""
    from pdds_amd.engine import default_run
    mesh, prob, opb = cfg2
    n = 2000
    print("cfg2 assembly, FP64, %d paths per knot" % n, flush=True)
    out = eng.assemble(mesh, prob, default_run(h=CFG2_H, seed=CFG2_SEED, sub_batch=32, item_paths=640), n)
    rng = np.random.default_rng(3)
    mc = np.flatnonzero(~mesh.on_boundary)
    knots = [int(k) for k in rng.choice(mc, 5, replace=False)] + list(CFG2_KNOTS.values())
    for k in knots:
        st, cs = _cshape(mesh, k)
        paths = O.simulate(opb, O.run_params(CFG2_H, CFG2_SEED, threads=8), mesh.x0[k], mesh.y0[k], st["rect"], k, 0, n)
        ref = O.reduce_batched(opb, paths, st, mesh.glob, cs, st["ipsi"], 32, 640, quirk=1)
        a, b = int(mesh.row_ptr[k]), int(mesh.row_ptr[k + 1])
        assert np.array_equal(mesh.cols[a:b], ref["cols"]), k
        assert np.array_equal(out["G"][a:b], ref["G"]), k
        assert np.array_equal(out["Gvar"][a:b], ref["Gvar"]), k
        assert out["B"][k] == ref["B"] and out["BB"][k] == ref["BB"], k


@pytest.mark.parametrize("which", list(CFG2_KNOTS))
def test_cfg2_knot_fp32_against_fp64_oracle(eng, cfg2, which):
    """FP32 path state (Kahan Z and xi) on the same Philox words as an FP64-stepped oracle: every
    G entry and B within half a standard error, and nearly every path of equal length."""
    from pdds_amd.engine import default_run
    mesh, prob, opb = cfg2
    k, n = CFG2_KNOTS[which], 20000
    st, cs = _cshape(mesh, k)
    print("cfg2 knot %d FP32" % k, flush=True)
    run = default_run(h=CFG2_H, seed=CFG2_SEED, precision="fp32", sub_batch=32, item_paths=640)
    got = eng.solve_knot(mesh, prob, run, k, n)
    paths = O.simulate(opb, O.run_params(CFG2_H, CFG2_SEED, normals=1, threads=8), mesh.x0[k], mesh.y0[k],
                       st["rect"], k, 0, n)
    ref = O.reduce_batched(opb, paths, st, mesh.glob, cs, st["ipsi"], 32, 640, quirk=1)
    se = np.sqrt(np.maximum(ref["Gvar"], 0.0) / n)
    assert np.all(np.abs(got["G"] - ref["G"]) <= 0.5 * se + 1e-12), np.max(np.abs(got["G"] - ref["G"]) / (se + 1e-300))
    seB = np.sqrt(max(ref["BB"] - ref["B"] ** 2, 0.0) / n)
    assert abs(got["B"] - ref["B"]) <= 0.5 * seB
    steps = int(paths["steps"].sum())
    assert abs(got["counters"]["path_steps"] - steps) <= 1e-4 * steps


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cfg2_whole_assembly_properties(eng, cfg2, precision):
    """Every knot of the benchmark mesh at 1e3 paths per knot: no unclassified exit, no capped
    path, the path-step count of the benchmark's scale, and the interface error of the bias floor
    (time step and RBF interpolation; it is the same from 1e3 to 1e5 paths per knot)."""
    from pdds_amd.engine import default_run
    mesh, prob, _ = cfg2
    print("cfg2 assembly %s" % precision, flush=True)
    out = eng.assemble(mesh, prob, default_run(h=CFG2_H, seed=CFG2_SEED, precision=precision), 1000)
    c = out["counters"]
    mc = int((~mesh.on_boundary).sum())
    assert mc == 3409 and c["paths"] == 1000 * mc
    assert c["exits"][5] == 0 and c["capped"] == 0
    assert 70 < c["path_steps"] / c["paths"] < 100
    u, it, rr = eng.solve(mesh.n_knots, mesh.row_ptr, mesh.cols, out["G"], out["B"], tol=1e-12)
    rms = float(np.sqrt(np.mean((u - prob.u(mesh.px, mesh.py)) ** 2)))
    assert 4.0e-3 <= rms <= 5.5e-3, rms
    assert np.all(np.isfinite(out["G"])) and np.all(np.isfinite(out["B"])) and np.all(out["Gvar"] >= 0.0)


# ------------------------------------------------------------------------------ serving --
DOMAIN = (-1.0, -1.0, 1.0, 1.0)
SRV_H, SRV_SEED, SRV_P = 2e-3, 23, 640


def _oracle_point(x, y, n, k2, stream):
    pb = O.problem(O.PDE_HELMHOLTZ, k2=k2)
    glob = np.array(DOMAIN)
    paths = O.simulate(pb, O.run_params(SRV_H, SRV_SEED, stream=stream, threads=8), x, y, glob, 0, 0, n)
    st = {"rect": glob, "msz": np.zeros(4, dtype=np.int32), "idx": np.zeros(0, dtype=np.int32),
          "px": np.zeros(0), "py": np.zeros(0)}
    return O.reduce_batched(pb, paths, st, glob, [1.0] * 4, [np.zeros((0, 0))] * 4, 32, SRV_P)


def test_serving_round_against_oracle(eng):
    """Points with their own streams and k^2 and a 4x4 x 16-knot solve served together (FP64
    paths): every point equals the oracle on its stream; the solve equals pdds_assemble +
    pdds_solve on the request's stream and k^2, two of whose rows equal the oracle's."""
    from pdds_amd.engine import Mesh, Server, bvp, default_run
    from pdds_amd.workloads import relative_fac
    rng = np.random.default_rng(29)
    pts = [(float(a), float(b), int(n), float(k) ** 2, 300 + i) for i, (a, b, n, k) in
           enumerate(zip(rng.uniform(-0.9, 0.9, 10), rng.uniform(-0.9, 0.9, 10), rng.integers(2000, 9000, 10),
                         rng.uniform(0.0, 4.0, 10)))]
    n_sub, n_node, n_solve, k2s, s_solve = 4, 16, 1000, 5.3, 77
    fac = relative_fac(DOMAIN[:2], DOMAIN[2:], n_sub, n_node, 0.8)
    run = default_run(h=SRV_H, seed=SRV_SEED, precision="fp64", sub_batch=32, item_paths=SRV_P)
    srv = Server(eng, bvp("helmholtz"), run, DOMAIN)
    tickets = [srv.submit_point(x, y, n, k2=k2, stream=s) for (x, y, n, k2, s) in pts[:5]]
    t_solve = srv.submit_solve(n_sub, n_node, fac, n_solve, k2=k2s, stream=s_solve, tol=1e-12)
    tickets += [srv.submit_point(x, y, n, k2=k2, stream=s) for (x, y, n, k2, s) in pts[5:]]
    res = [srv.wait(t) for t in tickets]
    sol = srv.wait(t_solve)
    srv.close()
    for (x, y, n, k2, s), r in zip(pts, res):
        ref = _oracle_point(x, y, n, k2, s)
        assert r["status"] == 0
        assert r["u"] == ref["B"] and r["u_var"] == ref["BB"] - ref["B"] ** 2, (x, y, k2, s)
    # the solve: served u == direct assembly + device solve, bit for bit
    m = Mesh(DOMAIN[:2], DOMAIN[2:], n_sub, n_node, fac)
    prob = bvp("helmholtz", k2=k2s)
    drun = default_run(h=SRV_H, seed=SRV_SEED, precision="fp64", sub_batch=32, item_paths=SRV_P, stream=s_solve)
    out = eng.assemble(m, prob, drun, n_solve)
    u, it, rr = eng.solve(m.n_knots, m.row_ptr, m.cols, out["G"], out["B"], tol=1e-12)
    assert sol["status"] == 0 and sol["n_knots"] == m.n_knots
    assert np.array_equal(sol["u_knots"], u) and sol["iters"] == it
    # two rows of that assembly against the oracle on the request's stream and k^2
    pb = O.problem(O.PDE_HELMHOLTZ, k2=k2s)
    mc = np.flatnonzero(~m.on_boundary)
    for k in (int(mc[len(mc) // 3]), int(mc[-7])):
        st = m.stencil(k)
        paths = O.simulate(pb, O.run_params(SRV_H, SRV_SEED, stream=s_solve, threads=8), m.x0[k], m.y0[k],
                           st["rect"], k, 0, n_solve)
        ref = O.reduce_batched(pb, paths, st, m.glob, st["cshape"], st["ipsi"], 32, SRV_P, quirk=1)
        sl = slice(m.row_ptr[k], m.row_ptr[k + 1])
        assert np.array_equal(out["G"][sl], ref["G"]) and out["B"][k] == ref["B"], k
    # and at the configuration's size the estimate is the manufactured solution's
    ex = prob.u(m.px, m.py)
    assert np.sqrt(np.mean((u - ex) ** 2)) < 0.05


# ------------------------------------------------------------------- raw-sample LUTs --
def _raw_pair(eng, samples, box):
    ny, nx, nf = samples.shape
    lut = eng.lut(samples, box, mode=1)
    sx, sy = nx / (box[2] - box[0]), ny / (box[3] - box[1])
    olut = {"coef": np.ascontiguousarray(samples, dtype=np.float32), "NX": nx, "NY": ny, "nf": nf,
            "ox": box[0] + 0.5 / sx, "oy": box[1] + 0.5 / sy, "sx": sx, "sy": sy}
    return lut, olut


def test_raw_lut_device_eval_matches_oracle(eng):
    rng = np.random.default_rng(12)
    box = (-1.0, -0.5, 1.0, 1.5)
    samples = rng.normal(0.0, 1.0, (7, 9, 5))
    lut, olut = _raw_pair(eng, samples, box)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut)
    px = rng.uniform(-1.4, 1.4, 3000)
    py = rng.uniform(-0.9, 1.9, 3000)
    for f in range(5):
        dev = lut.eval(f, px, py)
        ref = np.array([O.lib().orc_lut_eval(pb, f, a, b) for a, b in zip(px, py)])
        assert np.array_equal(dev, ref), f


@pytest.mark.parametrize("plain,lut_lds", [(0, 1), (0, 0), (1, 0)], ids=["tiled", "persistent", "plain"])
def test_raw_lut_paths_bit_exact(eng, plain, lut_lds):
    """The coefficient fields sampled at texel centres and used raw (mode 1) in every path kernel."""
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac, varcoef_fields
    n = 96
    box = (-1.0, -1.0, 1.0, 1.0)
    c = box[0] + (np.arange(n) + 0.5) * (box[2] - box[0]) / n
    X, Y = np.meshgrid(c, c)
    lut, olut = _raw_pair(eng, varcoef_fields(X, Y), box)
    fac = relative_fac(box[:2], box[2:], 4, 8, 1.0)
    m = Mesh(box[:2], box[2:], 4, 8, fac)
    k, npaths = 43, 3000
    st = m.stencil(k)
    run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, plain=plain, lut_lds=lut_lds)
    got = eng.solve_knot(m, bvp("varcoef_lut", lut=lut), run, k, npaths)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut)
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[k], m.y0[k], st["rect"], k, 0, npaths)
    ref = O.reduce_batched(pb, paths, st, m.glob, st["cshape"], st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
