"""Reflecting (Neumann / Robin) sides on the GPU: the Lepingle step in every kernel that takes it
(pdds_paths, the persistent and plain assembly kernels), FP64 bit-exact with the oracle's
restatement; FP32 within the Monte Carlo error of the FP64-stepped oracle on the same streams;
the unsupported combinations refused on the host."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GLOB = np.array([-1.0, -1.0, 1.0, 1.0])
PDE_ID = {"sinsin": O.PDE_SINSIN, "helmholtz": O.PDE_HELMHOLTZ, "varcoef": O.PDE_VARCOEF,
          "monegros": O.PDE_MONEGROS}


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


def _gmask(rect, glob=GLOB):
    """stencil sides on the global boundary: bit 0 N, 1 S, 2 E, 3 W"""
    return (int(rect[3] == glob[3]) | int(rect[1] == glob[1]) << 1 | int(rect[2] == glob[2]) << 2
            | int(rect[0] == glob[0]) << 3)


CASES = [
    # pde, k2, refl, robin, rho, rect
    ("sinsin", 0.0, 8, 0.0, 0.0, GLOB),
    ("helmholtz", 4.0, 10, -1.5, 0.0, GLOB),
    ("varcoef", 0.0, 3, 0.0, 0.05, GLOB),
    ("sinsin", 0.0, 12, -0.5, 0.0, np.array([-0.6, -0.9, 0.4, 0.2])),
]


@pytest.mark.parametrize("case", CASES, ids=["neumann-W", "robin-SW", "neumann-NS-band", "robin-EW-rect"])
def test_paths_fp64_bit_exact(eng, case):
    from pdds_amd.engine import bvp, default_run
    pde, k2, refl, robin, rho, rect = case
    x0, y0 = 0.5 * (rect[0] + rect[2]) - 0.05, 0.5 * (rect[1] + rect[3]) + 0.03
    n = 3000
    print("reflect paths %s refl=%d robin=%g rho=%g" % (pde, refl, robin, rho), flush=True)
    got = eng.paths(bvp(pde, k2=k2, neumann=refl, varphi=robin, rho=rho), default_run(h=2e-3, seed=7), x0, y0,
                    rect, 5, 0, n, phase2=True)
    pb = O.problem(PDE_ID[pde], k2=k2, refl=refl, robin=robin, rho=rho)
    ref = O.simulate(pb, O.run_params(2e-3, 7, phase2=1, threads=8), x0, y0, rect, 5, 0, n, phase2_out=True)
    for k in ("steps", "xt", "yt", "Y", "Z", "xi"):
        assert np.array_equal(got[k], ref[k]), (k, np.flatnonzero(got[k] != ref[k])[:5])
        assert np.array_equal(got["phase2"][k], ref["phase2"][k]), ("phase2", k)
    # reflecting sides are never exits (away from the corners)
    for s, (coord, val) in enumerate((("yt", rect[3]), ("yt", rect[1]), ("xt", rect[2]), ("xt", rect[0]))):
        if (refl >> s) & 1:
            other = "xt" if coord == "yt" else "yt"
            lo, hi = (rect[0], rect[2]) if other == "xt" else (rect[1], rect[3])
            on = (np.abs(got[coord] - val) < 1e-12) & (got[other] > lo + 1e-9) & (got[other] < hi - 1e-9)
            assert not on.any(), s


@pytest.mark.parametrize("plain", [0, 1], ids=["persistent", "plain"])
def test_point_eval_fp64_bit_exact(eng, plain):
    from pdds_amd.engine import bvp, default_run
    x = np.array([-0.8, -0.5, 0.2, 0.6])
    y = np.array([0.1, -0.4, 0.3, -0.85])
    n, k2, refl, robin = 4096, 2.0, 10, -1.0
    out = eng.point_eval(bvp("helmholtz", k2=k2, neumann=refl, varphi=robin), default_run(h=1e-3, seed=5, item_paths=512,
                                                                                         plain=plain), GLOB, x, y, n)
    pb = O.problem(O.PDE_HELMHOLTZ, k2=k2, refl=refl, robin=robin)
    st = {"rect": GLOB, "msz": np.zeros(4, dtype=np.int32), "idx": np.zeros(0, dtype=np.int32),
          "px": np.zeros(0), "py": np.zeros(0)}
    for i in range(len(x)):
        paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), x[i], y[i], GLOB, i, 0, n)
        ref = O.reduce_batched(pb, paths, st, GLOB, [1.0] * 4, [np.zeros((0, 0))] * 4, 32, 512)
        assert out["u"][i] == ref["B"], i
    u = bvp("helmholtz").u(x, y)
    assert np.all(np.abs(out["u"] - u) < 5 * np.sqrt(out["u_var"] / n) + 0.02)


@pytest.mark.parametrize("pde", ["sinsin", "lut"])
@pytest.mark.parametrize("plain", [0, 1], ids=["persistent", "plain"])
def test_knot_on_reflecting_side_fp64_bit_exact(eng, pde, plain):
    """A knot whose stencil's west side lies on the reflecting global west side: its G row, G
    variances, B and BB equal the oracle's with that stencil side reflecting."""
    from pdds_amd.engine import Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac, varcoef_fields
    fac = relative_fac((-1.0, -1.0), (1.0, 1.0), 4, 8, 1.0)
    m = Mesh((-1.0, -1.0), (1.0, 1.0), 4, 8, fac)
    k, n, refl, robin = 8, 3000, 8 | 2, -0.7
    st = m.stencil(k)
    assert st["rect"][0] == -1.0 and _gmask(st["rect"]) & 8
    if pde == "lut":
        nl = 96
        xs = np.linspace(-1, 1, nl)
        X, Y = np.meshgrid(xs, xs)
        samples = varcoef_fields(X, Y)
        lut = eng.lut(samples, (-1, -1, 1, 1), mode=0)
        d = 2.0 / (nl - 1)
        olut = {"coef": O.lut_prefilter(samples), "NX": nl + 2, "NY": nl + 2, "nf": 5, "ox": -1 - d, "oy": -1 - d,
                "sx": 1 / d, "sy": 1 / d}
        prob = bvp("varcoef_lut", lut=lut, neumann=refl, varphi=robin)
        pb = O.problem(O.PDE_VARCOEF_LUT, lut=olut, refl=_gmask(st["rect"]) & refl, robin=robin)
    else:
        prob = bvp("sinsin", neumann=refl, varphi=robin)
        pb = O.problem(O.PDE_SINSIN, refl=_gmask(st["rect"]) & refl, robin=robin)
    print("reflect knot %d %s plain=%d" % (k, pde, plain), flush=True)
    run = default_run(h=1e-3, seed=13, sub_batch=32, item_paths=640, plain=plain)
    got = eng.solve_knot(m, prob, run, k, n)
    paths = O.simulate(pb, O.run_params(1e-3, 13, threads=8), m.x0[k], m.y0[k], st["rect"], k, 0, n)
    ref = O.reduce_batched(pb, paths, st, m.glob, st["cshape"], st["ipsi"], 32, 640, quirk=1)
    assert np.array_equal(got["G"], ref["G"]) and np.array_equal(got["Gvar"], ref["Gvar"])
    assert got["B"] == ref["B"] and got["BB"] == ref["BB"]
    assert got["counters"]["path_steps"] == int(paths["steps"].sum())
    # no path leaves through the reflecting west side
    assert got["counters"]["exits"][3] == int(ref["cnt"][3]) and got["counters"]["exits"][5] == 0


def test_fp32_against_fp64_oracle(eng):
    from pdds_amd.engine import bvp, default_run
    x0, y0, n = -0.7, 0.2, 40000
    got = eng.paths(bvp("sinsin", neumann=8, varphi=-1.0), default_run(h=1e-3, seed=21, precision="fp32"), x0, y0,
                    GLOB, 2, 0, n)
    ref = O.simulate(O.problem(O.PDE_SINSIN, refl=8, robin=-1.0), O.run_params(1e-3, 21, normals=1, threads=8), x0, y0,
                     GLOB, 2, 0, n)
    same = got["steps"] == ref["steps"]
    assert same.mean() > 0.97, same.mean()

    def est(p):
        return p["Z"] + p["Y"] * (np.sin(np.pi * p["xt"]) * np.sin(np.pi * p["yt"]) + p["xt"] ** 2 + p["yt"] ** 2) + p["xi"]

    a, b = est(got), est(ref)
    assert abs(a.mean() - b.mean()) < 0.1 * b.std() / np.sqrt(n)
    assert abs(b.mean() - (np.sin(np.pi * x0) * np.sin(np.pi * y0) + x0 ** 2 + y0 ** 2)) < 4 * b.std() / np.sqrt(n) + 0.01


def test_unsupported_combinations_are_refused(eng):
    from pdds_amd._capi import ERR_INVALID, ERR_UNSUPPORTED, PddsError
    from pdds_amd.engine import Mesh, Server, bvp, default_run
    from pdds_amd.workloads import relative_fac
    m = Mesh((-1.0, -1.0), (1.0, 1.0), 2, 8, relative_fac((-1.0, -1.0), (1.0, 1.0), 2, 8, 1.0))
    with pytest.raises(PddsError) as ei:       # boundary knots on the reflecting side
        eng.assemble(m, bvp("sinsin", neumann=8), default_run(h=2e-3), 500)
    assert ei.value.status == ERR_UNSUPPORTED
    with pytest.raises(PddsError) as ei:
        eng.point_eval(bvp("sinsin", neumann=8, varphi=0.5), default_run(h=2e-3), GLOB, [0.1], [0.1], 100)
    assert ei.value.status == ERR_INVALID      # varphi must be <= 0
    with pytest.raises(PddsError) as ei:
        Server(eng, bvp("helmholtz", neumann=1), default_run(h=2e-3), tuple(GLOB))
    assert ei.value.status == ERR_UNSUPPORTED
    # the engine stays usable
    out = eng.point_eval(bvp("sinsin", neumann=8), default_run(h=2e-3), GLOB, [0.1], [0.1], 1000)
    assert np.isfinite(out["u"][0])
