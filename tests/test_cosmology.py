"""Host FP64 cosmology of libct2hip.so (cosmology.cpp:11-102 with explicit
parameters) against the CPU oracle and Einstein-de Sitter closed forms."""
import numpy as np
import pytest

from ct2_amd import api


@pytest.mark.parametrize("om,ol", [(1.0, 0.0), (0.3, 0.7), (0.25, 0.75)])
def test_matches_oracle(orc, om, ol):
    c = api.cosmology(2.0, om, 1e-4, ol - 1e-4)
    o = orc.cosmo(2.0, om, 1e-4, ol - 1e-4)
    for a in (0.02, 0.5, 1.0):
        assert api.cosmos_dadt(c, a) == pytest.approx(orc.cosmos_dadt(o, a), rel=1e-15)
    assert api.cosmos_time(c, 0.02, 0.5) == pytest.approx(orc.cosmos_time(o, 0.02, 0.5), rel=1e-14)
    assert api.cosmos_conformal_time(c, 0.02, 0.5) == pytest.approx(orc.cosmos_conformal_time(o, 0.02, 0.5), rel=1e-9)
    a = api.cosmos_update(c, 0.0, api.cosmos_dadt(c, 0.1), 0.1, 0.01)
    b = orc.cosmos_update(o, 0.0, orc.cosmos_dadt(o, 0.1), 0.1, 0.01)
    assert np.allclose(a, b, rtol=1e-14, atol=0)


def test_leapfrog_factors_closed_form():
    """EdS: int dt/a = (2/H)(a1^1/2 - a0^1/2), int dt/a^2 = (2/H)(a0^-1/2 - a1^-1/2)."""
    H = np.sqrt(8 * np.pi / 3)
    c = api.cosmology(H)
    for a0, a1 in ((0.02, 0.021), (0.02, 0.03), (0.5, 1.0)):
        assert api.cosmos_kick_factor(c, a0, a1) == pytest.approx(2 / H * (a1 ** 0.5 - a0 ** 0.5), rel=1e-11)
        assert api.cosmos_drift_factor(c, a0, a1) == pytest.approx(2 / H * (a0 ** -0.5 - a1 ** -0.5), rel=1e-11)
        assert api.cosmos_conformal_time(c, a0, a1) == pytest.approx(api.cosmos_kick_factor(c, a0, a1), rel=1e-8)
