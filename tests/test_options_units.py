"""Options block and code units (options.hpp:14-32, options.cpp:51-95) in
libct2hip.so: defaults, the derived fields against the oracle's restatement of
options.cpp:88-95, and against physics the formulas must satisfy (the box's
mean density is omega_m times the critical density; the radiation density of
a 2.73 K black body plus N_eff neutrinos against the textbook photon density;
the density parameters sum to one)."""
import math

import numpy as np
import pytest

from ct2_amd import api
from ct2_amd._lib import Ct2Error

G, C_LIGHT = 6.67259e-8, 2.99792458e10


def ulp_close(a, b, n=1):
    a, b = np.float32(a), np.float32(b)
    return abs(float(a) - float(b)) <= n * float(np.spacing(max(abs(a), abs(b))))


def test_defaults():
    o = api.options()
    assert (o.Ngrid, o.ppcell) == (32, 256)
    for k, v in (("hsoft", 1 / 50), ("bwidth", 3.0), ("omega_m", 0.25), ("mass_res", 1e9), ("hubble", 0.7),
                 ("eta", 0.15), ("Theta", 1.0)):
        assert getattr(o, k) == np.float32(v), k


@pytest.mark.parametrize("ngrid,mass_res,hubble,omega_m,theta",
                         [(32, 1e9, 0.7, 0.25, 1.0), (128, 2.5e10, 0.6774, 0.3089, 1.0), (256, 1e8, 0.5, 1.0, 0.9)])
def test_derived_fields_match_oracle(orc, ngrid, mass_res, hubble, omega_m, theta):
    o = api.options(Ngrid=ngrid, mass_res=mass_res, hubble=hubble, omega_m=omega_m, Theta=theta)
    w = orc.options_derive(ngrid, o.mass_res, o.hubble, o.omega_m, o.Theta)
    assert o.nparts == ngrid ** 3 and o.GM == w["GM"] == np.float32(1.0 / ngrid ** 3)
    for k in ("omega_r", "omega_lam"):
        assert ulp_close(getattr(o, k), w[k]), (k, getattr(o, k), w[k])
    for k in ("code_to_g", "code_to_cm", "code_to_s"):
        assert getattr(o, k) == pytest.approx(w[k], rel=1e-14), k
    c = api.options_cosmology(o)
    assert c.H == pytest.approx(w["H"], rel=1e-14)
    assert (c.omega_m, c.omega_r, c.omega_lam) == (o.omega_m, o.omega_r, o.omega_lam)


@pytest.mark.parametrize("source", ["library", "oracle"])
@pytest.mark.parametrize("ngrid,mass_res,hubble,omega_m,theta", [(32, 1e9, 0.7, 0.25, 1.0), (64, 3e10, 0.6, 0.3, 1.1)])
def test_physics_of_the_units(orc, source, ngrid, mass_res, hubble, omega_m, theta):
    if source == "library":
        o = api.options(Ngrid=ngrid, mass_res=mass_res, hubble=hubble, omega_m=omega_m, Theta=theta)
        f = {k: getattr(o, k) for k in ("GM", "code_to_g", "code_to_cm", "code_to_s", "omega_r", "omega_lam")}
        f["H"] = api.options_cosmology(o).H
    else:
        f = orc.options_derive(ngrid, mass_res, hubble, omega_m, theta)
    h = float(np.float32(hubble))
    om = float(np.float32(omega_m))
    H = h * 1e7 / 3.086e24  # 100 h km/s/Mpc in 1/s
    # mean density of ngrid^3 particles of mass_res suns in a box of side code_to_cm
    rho = ngrid ** 3 * 1.989e33 * float(np.float32(mass_res)) / f["code_to_cm"] ** 3
    assert rho == pytest.approx(om * 3 * H * H / (8 * math.pi * G), rel=1e-13)
    assert f["code_to_g"] == pytest.approx(1.989e33 * float(np.float32(mass_res)), rel=1e-15)
    assert f["code_to_s"] == pytest.approx(f["code_to_cm"] / C_LIGHT, rel=1e-15)
    # photons: omega_gamma h^2 = 2.4728e-5 at 2.7255 K (scaled as T^4); each of
    # 3.086 neutrino species adds 7/8 (4/11)^(4/3) of that
    T = 2.73 * float(np.float32(theta))
    og = 2.4728e-5 * (T / 2.7255) ** 4 / h ** 2
    assert float(f["omega_r"]) == pytest.approx(og * (1 + 3.086 * 7 / 8 * (4 / 11) ** (4 / 3)), rel=2e-3)
    assert float(f["omega_lam"]) + float(f["omega_r"]) + om == pytest.approx(1.0, abs=2e-7)
    # the Hubble rate in code time units
    assert f["H"] == pytest.approx(H * float(f["code_to_s"]), rel=1e-12)
    assert float(f["GM"]) * ngrid ** 3 == pytest.approx(1.0, rel=1e-7)


def test_reference_fp32_units_overflow_at_its_defaults():
    """Why the code units are FP64: stored in FP32 (options.hpp:17-19) the
    default mass unit, 1e9 solar masses, is out of range."""
    with np.errstate(over="ignore"):
        assert np.isinf(np.float32(1.989e33) * np.float32(1e9))
    o = api.options()
    assert math.isfinite(o.code_to_g) and math.isfinite(api.options_cosmology(o).H)


def test_cosmology_of_the_options_is_flat():
    o = api.options(Ngrid=64, omega_m=0.3, hubble=0.68)
    c = api.options_cosmology(o)
    # at a = 1 the expansion rate is H (omega_m + omega_r + omega_lam = 1)
    assert api.cosmos_dadt(c, 1.0) == pytest.approx(c.H, rel=1e-7)


@pytest.mark.parametrize("bad", [dict(Ngrid=0), dict(omega_m=0.0), dict(hubble=-1.0), dict(mass_res=0.0)])
def test_invalid_options_are_rejected(bad):
    with pytest.raises(Ct2Error):
        api.options(**bad)
