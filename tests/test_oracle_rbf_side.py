"""The reduction's Gaussian RBF values on a rectangle side (orc_rbf_side: one exponential at each
chunk's first entry, a two-multiply recurrence along the chunk), pinned against the direct
exp(-|p_j - X|^2 / c^2) of every entry (numpy, libm) on straight sides in both orientations, with
the exit inside the side, between two entries, on an entry, past either end, and at chunk
boundaries.  A wrong ratio (sign of d.s, a missing |s|^2 term, marching the wrong way) puts whole
chunks off by orders of magnitude and fails here."""
import numpy as np
import pytest

from oracle import oracle as O


def _side(m, horizontal, t0=-0.7, dt=0.0213, line=0.35):
    t = t0 + dt * np.arange(m)
    c = np.full(m, line)
    return (t, c) if horizontal else (c, t)


@pytest.mark.parametrize("m", [1, 2, 15, 16, 17, 33, 94])
@pytest.mark.parametrize("horizontal", [True, False], ids=["NS-side", "EW-side"])
def test_rbf_side_equals_direct(m, horizontal):
    px, py = _side(m, horizontal)
    width = 3.1 * 0.0213
    ic2 = 1.0 / (width * width)
    tpos = list(np.linspace(-0.8, -0.7 + 0.0213 * (m + 2), 23)) + [-0.7 + 0.0213 * (16 + 0.5), -0.7 + 0.0213 * 16]
    for t in tpos:
        x, y = (t, 0.35) if horizontal else (0.35, t)
        got = O.rbf_side(px, py, x, y, ic2)
        arg = ((px - x) ** 2 + (py - y) ** 2) * ic2
        want = np.exp(-arg)
        big = want > 1e-180
        # the recurrence's rounding grows as k^2 ulp along a chunk; the direct form's own error
        # grows with the exponent's size (its argument carries a relative rounding error)
        tol = (1e-13 + 1e-15 * arg[big]) * want[big]
        assert np.all(np.abs(got[big] - want[big]) <= tol), (m, t)
        # a chunk whose first entry underflows gives 0 for values the direct form takes below
        # about 1e-180 (nothing a sum of Y phi over paths can see)
        assert np.all(got[~big] < 1e-170)


def test_rbf_side_exit_off_the_line():
    """The exit's distance from the side's line enters every value through d, like the direct
    form (the engine's exits lie on their side; the form does not depend on that)."""
    px, py = _side(40, True)
    ic2 = 1.0 / (0.05 ** 2)
    got = O.rbf_side(px, py, -0.31, 0.37, ic2)
    want = np.exp(-((px + 0.31) ** 2 + (py - 0.37) ** 2) * ic2)
    big = want > 1e-180
    assert np.all(np.abs(got[big] - want[big]) <= 1e-13 * want[big])
