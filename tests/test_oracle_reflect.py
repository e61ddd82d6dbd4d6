"""The oracle's reflected (Neumann / Robin) paths, pinned by the mathematics.

The Lepingle step (Integrators/FKACIntegrator.hpp:114-165) reflects paths off the sides in the
problem's mask; a local time L accrues there, Z gains psi Y L and Y the factor of varphi L.  On
[-1,1]^2 with u = sin(pi x) sin(pi y) + x^2 + y^2, sigma = sqrt(2) I, psi = du/dn - varphi u on
the reflecting sides and g = u on the others, the Feynman-Kac estimate E[Z + Y g(X_tau) + xi]
equals u(x0) for any varphi <= 0 (Ito's formula for the reflected diffusion); so the estimates
must converge to u as h -> 0, at first order.  A mistake in the local time (its sign, a factor,
the bridge maximum, where psi is taken) leaves an O(1) or O(sqrt h) error and fails here.
"""
import numpy as np
import pytest

from oracle import oracle as O

GLOB = np.array([-1.0, -1.0, 1.0, 1.0])
PTS = [(-0.8, 0.1), (-0.5, -0.4), (0.2, 0.3)]


def u_exact(x, y):
    return np.sin(np.pi * x) * np.sin(np.pi * y) + x * x + y * y


def _errors(refl, robin, h, n=24000, seed=5):
    out = []
    for i, (x, y) in enumerate(PTS):
        pb = O.problem(O.PDE_SINSIN, refl=refl, robin=robin)
        p = O.simulate(pb, O.run_params(h, seed, threads=8), x, y, GLOB, i, 0, n)
        est = p["Z"] + p["Y"] * u_exact(p["xt"], p["yt"]) + p["xi"]
        out.append((est.mean() - u_exact(x, y), est.std() / np.sqrt(n), p))
    return out


@pytest.mark.parametrize("refl,robin", [(8, 0.0), (8, -2.0), (10, 0.0)], ids=["neumann-W", "robin-W", "neumann-SW"])
def test_reflected_estimate_converges_at_first_order(refl, robin):
    coarse = _errors(refl, robin, 4e-3)
    fine = _errors(refl, robin, 1e-3)
    e0 = max(abs(e) for e, _, _ in coarse)
    e1 = max(abs(e) for e, _, _ in fine)
    # first order: a fourth of h cuts the largest error about fourfold (noise 0.0015-0.003)
    assert e1 < 0.45 * e0 + 0.004, (e0, e1)
    assert e1 < 0.012, e1
    for e, se, p in fine:
        # no exit through a reflecting side (away from the corners)
        if refl & 8:
            assert not np.any((np.abs(p["xt"] + 1.0) < 1e-12) & (np.abs(p["yt"]) < 0.999))
        if refl & 2:
            assert not np.any((np.abs(p["yt"] + 1.0) < 1e-12) & (np.abs(p["xt"]) < 0.999))


def test_mask_zero_is_the_dirichlet_path():
    x, y = PTS[1]
    a = O.simulate(O.problem(O.PDE_HELMHOLTZ, k2=2.0), O.run_params(2e-3, 9, threads=4), x, y, GLOB, 3, 0, 500)
    b = O.simulate(O.problem(O.PDE_HELMHOLTZ, k2=2.0, refl=0, robin=-3.0, rho=0.05), O.run_params(2e-3, 9, threads=4),
                   x, y, GLOB, 3, 0, 500)
    for k in ("xt", "yt", "Y", "Z", "xi", "steps"):
        assert np.array_equal(a[k], b[k]), k


def test_rho_band_changes_paths_not_the_limit():
    """rho restricts the Lepingle step to a band; outside it overshoots are projected back.  Both
    estimate u (the band only changes how the local time is sampled)."""
    x, y = PTS[0]
    n, h = 24000, 1e-3
    ests = []
    for rho in (0.0, 0.05):
        p = O.simulate(O.problem(O.PDE_SINSIN, refl=8, rho=rho), O.run_params(h, 5, threads=8), x, y, GLOB, 0, 0, n)
        est = p["Z"] + p["Y"] * u_exact(p["xt"], p["yt"]) + p["xi"]
        ests.append((est.mean(), est.std() / np.sqrt(n)))
    for m, se in ests:
        assert abs(m - u_exact(x, y)) < 4 * se + 0.01
