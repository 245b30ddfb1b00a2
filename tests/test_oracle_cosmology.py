"""Pins the FP64 cosmology oracle (cosmology.cpp:11-102) with the closed
forms of an Einstein-de Sitter universe (Omega_m = 1): a ~ t^(2/3),
t = (2/3H) a^(3/2), conformal tau = int dt/a = (2/H) a^(1/2)."""
import numpy as np
import pytest

H = 2.0


def test_dadt_closed_form(orc):
    c = orc.cosmo(H)
    for a in (0.02, 0.1, 1.0):
        assert orc.cosmos_dadt(c, a) == pytest.approx(H * a ** -0.5, rel=1e-15)


def test_time_and_conformal_time(orc):
    c = orc.cosmo(H)
    a0, a1 = 0.02, 0.03
    assert orc.cosmos_time(c, a0, a1) == pytest.approx(2 / (3 * H) * (a1 ** 1.5 - a0 ** 1.5), rel=1e-10)
    assert orc.cosmos_conformal_time(c, a0, a1) == pytest.approx(2 / H * (a1 ** 0.5 - a0 ** 0.5), rel=1e-8)


def test_update_follows_conformal_solution(orc):
    """cosmos_update advances (a, adot) by conformal time dt0 (da = adot a dt):
    a(tau) = (sqrt(a0) + H tau/2)^2 and adot = H a^-1/2 for EdS."""
    c = orc.cosmo(H)
    a0 = 0.02
    errs = []
    for dtau in (0.05, 0.025):
        _, adot, a = orc.cosmos_update(c, 0.0, H * a0 ** -0.5, a0, dtau)
        a_exact = (np.sqrt(a0) + H * dtau / 2) ** 2
        assert a == pytest.approx(a_exact, rel=5e-8)
        assert adot == pytest.approx(H * a_exact ** -0.5, rel=5e-8)
        errs.append(abs(a / a_exact - 1))
    # three-stage third-order scheme with 128 fixed sub-steps: the global
    # error scales as dtau^4 / 128^3, so halving dtau divides it by ~16
    assert 12 < errs[0] / errs[1] < 20


def test_lambda_dominated_limit(orc):
    """Omega_L = 1: a = e^{H t}, so int dt = ln(a1/a0)/H."""
    c = orc.cosmo(H, omega_m=0.0, omega_lam=1.0)
    assert orc.cosmos_time(c, 0.5, 1.0) == pytest.approx(np.log(2.0) / H, rel=1e-10)
