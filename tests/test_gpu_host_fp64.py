"""The library's host FP64 parts on the GPU box, where the round-end run loads
it: the cosmology (cosmology.cpp:11-102) and the options block with its code
units (options.cpp:51-95) against the oracle, after the device is initialised
(the same calls the stepping loop makes between kicks)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("om,ol", [(1.0, 0.0), (0.3, 0.7)])
def test_cosmology_matches_oracle_on_the_gpu_box(ct2, orc, om, ol):
    c = ct2.cosmology(2.0, om, 1e-4, ol - 1e-4)
    o = orc.cosmo(2.0, om, 1e-4, ol - 1e-4)
    for a in (0.02, 0.5, 1.0):
        assert ct2.cosmos_dadt(c, a) == pytest.approx(orc.cosmos_dadt(o, a), rel=1e-15)
    assert ct2.cosmos_time(c, 0.02, 0.5) == pytest.approx(orc.cosmos_time(o, 0.02, 0.5), rel=1e-14)
    a = ct2.cosmos_update(c, 0.0, ct2.cosmos_dadt(c, 0.1), 0.1, 0.01)
    b = orc.cosmos_update(o, 0.0, orc.cosmos_dadt(o, 0.1), 0.1, 0.01)
    assert np.allclose(a, b, rtol=1e-14, atol=0)


def test_leapfrog_factors_on_the_gpu_box(ct2):
    """Einstein-de Sitter: int dt/a = (2/H)(a1^1/2 - a0^1/2), int dt/a^2 = (2/H)(a0^-1/2 - a1^-1/2)."""
    H = np.sqrt(8 * np.pi / 3)
    c = ct2.cosmology(H)
    a0, a1 = 0.02, 0.03
    assert ct2.cosmos_kick_factor(c, a0, a1) == pytest.approx(2 / H * (a1 ** 0.5 - a0 ** 0.5), rel=1e-11)
    assert ct2.cosmos_drift_factor(c, a0, a1) == pytest.approx(2 / H * (a0 ** -0.5 - a1 ** -0.5), rel=1e-11)


def test_options_match_oracle_on_the_gpu_box(ct2, orc):
    o = ct2.options(Ngrid=128, mass_res=2.5e10, hubble=0.6774, omega_m=0.3089)
    w = orc.options_derive(128, o.mass_res, o.hubble, o.omega_m, o.Theta)
    assert o.GM == w["GM"] == np.float32(1.0 / 128 ** 3)
    for k in ("code_to_g", "code_to_cm", "code_to_s"):
        assert getattr(o, k) == pytest.approx(w[k], rel=1e-14), k
    assert ct2.options_cosmology(o).H == pytest.approx(w["H"], rel=1e-14)
