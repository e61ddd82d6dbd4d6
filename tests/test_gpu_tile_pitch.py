"""The LDS tile's row pitch (run.lut_pad, the autotuner's tile-shape knob) moves which LDS banks
each tile row starts on and nothing else: a knot's G row, variances, B and BB are bit-identical for
every pitch, and equal to the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_tile_pitch_does_not_change_results():
    from pdds_amd.engine import Engine, Mesh, bvp, default_run
    from pdds_amd.workloads import relative_fac, varcoef_fields
    eng = Engine(0)
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    fac = relative_fac(sw, ne, 4, 8, 1.0)
    m = Mesh(sw, ne, 4, 8, fac)
    n = 96
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    samples = varcoef_fields(X, Y)
    prob = bvp("varcoef_lut", lut=eng.lut(samples, (-1, -1, 1, 1), mode=0))
    knot, npaths = 43, 3000
    out = {}
    for pad in (0, 1, 3, 7):
        for prec in ("fp64", "fp32"):
            print("tile pitch pad %d %s" % (pad, prec), flush=True)
            run = default_run(h=1e-3, seed=5, sub_batch=32, item_paths=640, lut_lds=1, lut_pad=pad, precision=prec)
            out[(pad, prec)] = eng.solve_knot(m, prob, run, knot, npaths)
    for pad in (1, 3, 7):
        for prec in ("fp64", "fp32"):
            a, b = out[(0, prec)], out[(pad, prec)]
            assert np.array_equal(a["G"], b["G"]) and np.array_equal(a["Gvar"], b["Gvar"]), (pad, prec)
            assert a["B"] == b["B"] and a["BB"] == b["BB"], (pad, prec)
    st = m.stencil(knot)
    pb = O.problem(O.PDE_VARCOEF_LUT, lut=O.lut_dict(samples, (-1, -1, 1, 1)))
    paths = O.simulate(pb, O.run_params(1e-3, 5, threads=8), m.x0[knot], m.y0[knot], st["rect"], knot, 0, npaths)
    ref = O.reduce_batched(pb, paths, st, m.glob, [fac[min(max(int(k), 0), 3)] for k in st["kind"]],
                           st["ipsi"], 32, 640, quirk=1)
    got = out[(3, "fp64")]
    assert np.array_equal(got["G"], ref["G"]) and got["B"] == ref["B"] and got["BB"] == ref["BB"]
    eng.close()
