"""The oracle's elementary pieces against published values and libm/mpmath."""
import math

import numpy as np
import pytest

from oracle import oracle as O


# Published Philox4x32-10 known-answer vectors (Salmon et al., SC'11 / Random123 kat_vectors)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(ctr, key, out):
    assert tuple(int(v) for v in O.philox(ctr, key)) == out


def ulp_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_exp_log_expm1_tanh_ulps():
    rng = np.random.default_rng(1)
    x = rng.uniform(-700, 700, 200000)
    assert ulp_err(O.math_eval("exp", x), np.exp(x)).max() <= 2
    y = np.exp(rng.uniform(-700, 700, 200000))
    assert ulp_err(O.math_eval("log", y), np.log(y)).max() <= 4
    z = rng.uniform(-0.5, 0.5, 200000)
    assert ulp_err(O.math_eval("expm1", z), np.expm1(z)).max() <= 3
    t = rng.uniform(-5, 5, 200000)
    assert ulp_err(O.math_eval("tanh", t), np.tanh(t)).max() <= 6
    # subnormal result region and the exact special values
    assert O.math_eval("exp", np.array([0.0]))[0] == 1.0
    assert O.math_eval("log", np.array([1.0]))[0] == 0.0
    assert abs(O.math_eval("exp", np.array([-744.0]))[0] - math.exp(-744.0)) <= 5e-324


def test_table_exp_ulps_against_mpmath():
    """The RBF reduction's table-driven exp (2^(j/64) table, degree-5 polynomial) on x <= 0:
    within 2 ulp of the correctly rounded exp, exact at 0, flushed to 0 below -708 (where the
    result would leave the normal range).  A wrong table entry or coefficient fails."""
    import mpmath as mp
    mp.mp.prec = 120
    rng = np.random.default_rng(4)
    x = np.concatenate([-rng.exponential(2.0, 3000), rng.uniform(-700, 0, 3000), -np.arange(64) * math.log(2) / 64])
    ref = np.array([float(mp.exp(mp.mpf(v))) for v in x])
    got = O.math_eval("exp_tab", x)
    assert ulp_err(got, ref).max() <= 2
    assert O.math_eval("exp_tab", np.array([0.0]))[0] == 1.0
    # the edge of the normal range, and the flush to zero below it
    edge = np.array([-708.0, -707.99, -707.5])
    assert ulp_err(O.math_eval("exp_tab", edge), np.array([float(mp.exp(mp.mpf(v))) for v in edge])).max() <= 2
    assert np.all(O.math_eval("exp_tab", np.array([-708.0001, -710.0, -746.0, -1e300])) == 0.0)
    # arguments where x * 64/ln2 lies on a half-integer, the rounding boundary of the reduction
    ties = (np.arange(-2001, 0, 2) + 0.5) * (math.log(2) / 64)
    assert ulp_err(O.math_eval("exp_tab", ties), np.array([float(mp.exp(mp.mpf(v))) for v in ties])).max() <= 2


def test_sinpi_cospi_and_sin_cos():
    rng = np.random.default_rng(2)
    x = rng.uniform(-50, 50, 200000)
    # exact reduction: sinpi at integers and half-integers is exact
    ints = np.arange(-20, 21, dtype=np.float64)
    assert np.all(O.math_eval("sinpi", ints) == 0.0)
    assert np.all(np.abs(O.math_eval("cospi", ints)) == 1.0)
    assert np.all(np.abs(O.math_eval("sinpi", ints + 0.5)) == 1.0)
    import mpmath as mp
    mp.mp.prec = 120
    xs = x[:2000]
    ref_s = np.array([float(mp.sinpi(mp.mpf(v))) for v in xs])
    ref_c = np.array([float(mp.cospi(mp.mpf(v))) for v in xs])
    assert np.abs(O.math_eval("sinpi", xs) - ref_s).max() < 4e-16
    assert np.abs(O.math_eval("cospi", xs) - ref_c).max() < 4e-16
    # sin/cos of general arguments: |x| * eps accuracy from forming x/pi
    assert np.abs(O.math_eval("sin", x) - np.sin(x)).max() < 1e-14
    assert np.abs(O.math_eval("cos", x) - np.cos(x)).max() < 1e-14


def test_normals_moments_and_independence():
    run = O.run_params(1e-3, 7)
    n1, n2 = O.normals(run, 3, 11, 0, 200000)
    for n in (n1, n2):
        assert abs(n.mean()) < 0.01
        assert abs(n.var() - 1) < 0.01
        assert abs(((n ** 4).mean()) - 3) < 0.06
    assert abs(np.corrcoef(n1, n2)[0, 1]) < 0.01
    # 24-bit (FP32 engine) streams: same statistics, consecutive steps from one draw
    run32 = O.run_params(1e-3, 7, normals=1)
    m1, m2 = O.normals(run32, 3, 11, 0, 200000)
    assert abs(m1.mean()) < 0.01 and abs(m1.var() - 1) < 0.01
    assert abs(np.corrcoef(m1[:-1], m1[1:])[0, 1]) < 0.01
    # keying: a different path or knot gives a different stream
    a1, _ = O.normals(run, 3, 12, 0, 1000)
    b1, _ = O.normals(run, 4, 11, 0, 1000)
    assert not np.allclose(a1, n1[:1000]) and not np.allclose(b1, n1[:1000])


def test_gobet_constant_literal():
    # the reference uses -0.5826 (Solvers/EMFKACSolver.hpp:186) for -zeta(1/2)/sqrt(2 pi)
    import mpmath as mp
    assert abs(float(-mp.zeta(0.5) / mp.sqrt(2 * mp.pi)) - 0.5826) < 1e-4
