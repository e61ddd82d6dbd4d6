"""The oracle's bicubic B-spline lookup (orc_lut_eval, orc_lut_prefilter), pinned independently.

Mode 1 is the reference's raw-sample lookup cubicTex2DSimple (Integrators/cubicTex2D.cu:64-87):
texel coordinate x = Nx (X - x0) / (x1 - x0) (Integrators/FKACIntegrator.cuh:32-33), shifted by
half a texel, 16 taps weighted by the uniform cubic B-spline bspline(t)
(Integrators/bspline_kernel.cu:52-60), clamp-to-edge addressing (BVPs/LUT2.cuh:35-41).  Pins:
a brute-force 16-tap sum written from those formulas, partition of unity, and exact reproduction
of linear data (the cubic B-spline reproduces polynomials of degree 1).

Mode 0 is the interpolating spline (prefiltered coefficients): it reproduces the samples at the
nodes and converges to a smooth function at fourth order between them.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _ref_bspline(t):
    """bspline(t) of Integrators/bspline_kernel.cu:52-60, in FP64."""
    t = abs(t)
    a = 2.0 - t
    if t < 1.0:
        return 2.0 / 3.0 - 0.5 * t * t * a
    if t < 2.0:
        return a * a * a / 6.0
    return 0.0


def _brute_simple(T, box, X, Y):
    """cubicTex2DSimple on raw samples T[ny][nx] over box, brute force in FP64."""
    ny, nx = T.shape
    x = nx * (X - box[0]) / (box[2] - box[0]) - 0.5
    y = ny * (Y - box[1]) / (box[3] - box[1]) - 0.5
    i, j = int(np.floor(x)), int(np.floor(y))
    fx, fy = x - i, y - j
    acc = 0.0
    for l in range(-1, 3):
        jj = min(max(j + l, 0), ny - 1)
        for k in range(-1, 3):
            ii = min(max(i + k, 0), nx - 1)
            acc += _ref_bspline(l - fy) * _ref_bspline(k - fx) * T[jj, ii]
    return acc


def _raw_problem(T, box):
    """The oracle's problem with a one-field mode-1 table: samples used raw as coefficients, the
    texel-centre map t = (X - ox) sx with sx = nx / (x1 - x0) and ox = x0 + 1 / (2 sx)."""
    ny, nx = T.shape
    sx, sy = nx / (box[2] - box[0]), ny / (box[3] - box[1])
    lut = {"coef": np.ascontiguousarray(T[:, :, None], dtype=np.float32), "NX": nx, "NY": ny, "nf": 1,
           "ox": box[0] + 0.5 / sx, "oy": box[1] + 0.5 / sy, "sx": sx, "sy": sy}
    return O.problem(O.PDE_VARCOEF_LUT, lut=lut)


def _ev(pb, x, y, f=0):
    return O.lib().orc_lut_eval(pb, f, float(x), float(y))


BOX = (-1.0, -0.5, 1.0, 1.5)


def test_raw_mode_matches_brute_force_16_tap_sum():
    rng = np.random.default_rng(3)
    T = rng.normal(0.0, 1.0, (7, 9)).astype(np.float32).astype(np.float64)
    pb = _raw_problem(T, BOX)
    # points inside, on texel centres and edges, and outside the box (clamped taps)
    xs = np.concatenate([rng.uniform(-1.3, 1.3, 400), np.linspace(-1, 1, 19)])
    ys = np.concatenate([rng.uniform(-0.8, 1.8, 400), np.linspace(-0.5, 1.5, 19)])
    for x, y in zip(xs, ys):
        ref = _brute_simple(T, BOX, x, y)
        assert abs(_ev(pb, x, y) - ref) <= 4e-6 * (1.0 + abs(ref)), (x, y)


def test_raw_mode_partition_of_unity():
    T = np.full((6, 5), 2.75)
    pb = _raw_problem(T, BOX)
    rng = np.random.default_rng(4)
    for x, y in zip(rng.uniform(-1.5, 1.5, 300), rng.uniform(-1.0, 2.0, 300)):
        assert abs(_ev(pb, x, y) - 2.75) <= 2.75 * 4e-7


def test_raw_mode_reproduces_linear_data():
    """Samples of a X + b Y + d at the texel centres come back exactly wherever every tap lies
    inside the table (two texels from the edge)."""
    ny, nx = 12, 16
    sx, sy = nx / (BOX[2] - BOX[0]), ny / (BOX[3] - BOX[1])
    xc = BOX[0] + (np.arange(nx) + 0.5) / sx
    yc = BOX[1] + (np.arange(ny) + 0.5) / sy
    XC, YC = np.meshgrid(xc, yc)
    T = 0.7 * XC - 1.3 * YC + 0.25
    pb = _raw_problem(T, BOX)
    rng = np.random.default_rng(5)
    xs = rng.uniform(xc[1], xc[-3], 300)
    ys = rng.uniform(yc[1], yc[-3], 300)
    for x, y in zip(xs, ys):
        assert abs(_ev(pb, x, y) - (0.7 * x - 1.3 * y + 0.25)) < 2e-6


def test_raw_mode_is_not_interpolating():
    """Raw samples are smoothed: at a texel centre of curved data the value differs from the
    sample (by (T[i-1] - 2 T[i] + T[i+1]) / 6 in 1-D), as the reference's lookup does."""
    T = np.zeros((8, 8))
    T[:, 4] = 6.0
    pb = _raw_problem(T, (0.0, 0.0, 8.0, 8.0))
    assert abs(_ev(pb, 4.5, 4.5) - 4.0) < 1e-6          # 6 * 2/3
    assert abs(_ev(pb, 3.5, 4.5) - 1.0) < 1e-6          # 6 * 1/6


def _node_problem(samples, box):
    """Mode 0: prefiltered coefficients on a (ny + 2) x (nx + 2) grid, node-aligned map."""
    ny, nx = samples.shape[:2]
    s = samples if samples.ndim == 3 else samples[:, :, None]
    dx, dy = (box[2] - box[0]) / (nx - 1), (box[3] - box[1]) / (ny - 1)
    lut = {"coef": O.lut_prefilter(s), "NX": nx + 2, "NY": ny + 2, "nf": s.shape[2], "ox": box[0] - dx,
           "oy": box[1] - dy, "sx": 1 / dx, "sy": 1 / dy}
    return O.problem(O.PDE_VARCOEF_LUT, lut=lut)


def _smooth(x, y):
    return np.sin(1.3 * x + 0.4) * np.cos(0.9 * y) + 0.2 * x * y


@pytest.mark.parametrize("n", [33, 65])
def test_prefiltered_mode_interpolates_and_converges(n):
    xs = np.linspace(BOX[0], BOX[2], n)
    ys = np.linspace(BOX[1], BOX[3], n)
    X, Y = np.meshgrid(xs, ys)
    pb = _node_problem(_smooth(X, Y), BOX)
    # the nodes come back (FP32 coefficients)
    for j in range(0, n, 5):
        for i in range(0, n, 3):
            assert abs(_ev(pb, xs[i], ys[j]) - _smooth(xs[i], ys[j])) < 1e-6
    # between the nodes: fourth-order convergence in the interior
    rng = np.random.default_rng(6)
    h = xs[1] - xs[0]
    px = rng.uniform(xs[4], xs[-5], 400)
    py = rng.uniform(ys[4], ys[-5], 400)
    err = max(abs(_ev(pb, a, b) - _smooth(a, b)) for a, b in zip(px, py))
    assert err < 0.05 * h ** 4 + 1e-6, (err, h)
    # and near the edges still second order at worst
    px = rng.uniform(BOX[0], BOX[2], 400)
    py = rng.uniform(BOX[1], BOX[3], 400)
    err = max(abs(_ev(pb, a, b) - _smooth(a, b)) for a, b in zip(px, py))
    assert err < 0.5 * h ** 2, (err, h)


def test_prefiltered_mode_fourth_order_rate():
    """Halving the node spacing cuts the interior error about 16x (an end condition or a
    prefilter term off would show a lower rate)."""
    def rough(x, y):   # large fourth derivatives keep the error above the FP32 rounding floor
        return np.sin(3.1 * x + 0.4) * np.cos(2.7 * y) + 0.2 * x * y

    errs = []
    rng = np.random.default_rng(8)
    px = rng.uniform(-0.5, 0.5, 300)
    py = rng.uniform(0.0, 1.0, 300)
    for n in (17, 33):
        xs = np.linspace(BOX[0], BOX[2], n)
        ys = np.linspace(BOX[1], BOX[3], n)
        X, Y = np.meshgrid(xs, ys)
        pb = _node_problem(rough(X, Y), BOX)
        errs.append(max(abs(_ev(pb, a, b) - rough(a, b)) for a, b in zip(px, py)))
    assert 12.0 < errs[0] / errs[1] < 20.0, errs
