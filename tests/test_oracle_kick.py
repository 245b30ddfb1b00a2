"""Pins the FP64 kick, epilogue and drift oracle with closed forms,
invariants and exact arithmetic (kick.cpp:36-99, drift.cpp:4-17)."""
from fractions import Fraction

import numpy as np
import pytest

TWO32 = 2 ** 32


def fx(v):
    """box coordinate -> fixed32 (exact for dyadic inputs used here)."""
    return np.array([int(round((a % 1.0) * TWO32)) % TWO32 for a in np.atleast_1d(v)], dtype=np.uint32)


def two_body(d, soft):
    x = fx([0.3, 0.3 + d])
    y = fx([0.5, 0.5])
    z = fx([0.25, 0.25])
    return x, y, z


@pytest.mark.parametrize("d", [0.001, 0.01, 0.2])
def test_two_body_plummer_closed_form(orc, d):
    soft, gm = 1.0 / 1600, 0.5
    x, y, z = two_body(d, soft)
    dd = (int(x[1]) - int(x[0])) / TWO32
    ax, ay, az, pot = orc.kick(x, y, z, gm, soft, orc.KERNEL_PLUMMER)
    f = gm * dd / (dd * dd + soft * soft) ** 1.5
    assert ax[0] == pytest.approx(f, rel=1e-14) and ax[1] == pytest.approx(-f, rel=1e-14)
    assert ay[0] == 0 and az[0] == 0
    p = gm * (1.0 / soft + 1.0 / np.sqrt(dd * dd + soft * soft))  # self pair included
    assert pot[0] == pytest.approx(p, rel=1e-14) and pot[1] == pytest.approx(p, rel=1e-14)


def test_two_body_ct2_kernel_both_branches(orc):
    """kick.cpp:52-66: inside h F = dx/h^3, pot = 1.5/h - r^2/(2h^3); outside
    Newtonian; the self pair adds 1.5/h to the potential only."""
    h, gm = 0.02, 1.0
    for d, near in ((0.005, True), (0.05, False)):
        x, y, z = two_body(d, h)
        dd = (int(x[1]) - int(x[0])) / TWO32
        ax, _, _, pot = orc.kick(x, y, z, gm, h, orc.KERNEL_CT2)
        if near:
            f, p = dd / h ** 3, 1.5 / h - 0.5 * dd * dd / h ** 3
        else:
            f, p = 1.0 / dd ** 2, 1.0 / dd
        assert ax[0] == pytest.approx(f, rel=1e-13)
        assert pot[0] == pytest.approx(1.5 / h + p, rel=1e-13)


def test_ct2_kernel_continuous_at_h(orc):
    h = 0.02
    a, b = [], []
    for d in (h * (1 - 1e-7), h * (1 + 1e-7)):
        x, y, z = two_body(d, h)
        r = orc.kick(x, y, z, 1.0, h, orc.KERNEL_CT2)
        a.append(r[0][0])
        b.append(r[3][0])
    assert a[0] == pytest.approx(a[1], rel=1e-6)
    assert b[0] == pytest.approx(b[1], rel=1e-6)


def test_minimum_image_across_boundary(orc):
    """Sources across the periodic boundary pull through it."""
    x = fx([0.01, 0.99])
    y = fx([0.5, 0.5])
    z = fx([0.5, 0.5])
    ax, _, _, _ = orc.kick(x, y, z, 1.0, 1e-3, orc.KERNEL_PLUMMER)
    d = 0.02
    assert ax[0] == pytest.approx(-d / (d * d + 1e-6) ** 1.5, rel=1e-6)
    assert ax[1] == pytest.approx(+d / (d * d + 1e-6) ** 1.5, rel=1e-6)


@pytest.mark.parametrize("kernel", [0, 1])
def test_newton_third_law(orc, kernel):
    rng = np.random.default_rng(11)
    n = 300
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    ax, ay, az, _ = orc.kick(x, y, z, 1.0 / n, 0.01, kernel)
    for a in (ax, ay, az):
        assert abs(a.sum()) <= 1e-12 * np.abs(a).sum()


def test_lattice_force_vanishes(orc):
    """Odd lattice: every source at +k has a partner at -k (no pair sits at
    exactly half a box, where the int32 minimum image is one-sided)."""
    n = 5
    g = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = fx(X.ravel()), fx(Y.ravel()), fx(Z.ravel())
    ax, ay, az, pot = orc.kick(x, y, z, 1.0 / n ** 3, 0.01, orc.KERNEL_PLUMMER)
    # (i+1/2)/5 is not dyadic: fixed32 rounding moves sites by <= 2^-33, which
    # leaves forces of order 1e-9 against a nearest-neighbour scale of n^2
    scale = 1.0 / (1.0 / n) ** 2
    assert np.max(np.abs([ax, ay, az])) < 1e-9 * scale
    assert np.ptp(pot) < 1e-9 * pot.mean()
    # a single displaced site breaks the symmetry by far more
    x2 = x.copy()
    x2[0] += np.uint32(2 ** 22)
    ax2 = orc.kick(x2, y, z, 1.0 / n ** 3, 0.01, orc.KERNEL_PLUMMER)[0]
    assert np.max(np.abs(ax2)) > 1e-5 * scale


@pytest.mark.parametrize("kernel", [0, 1])
def test_fp64_sum_matches_long_double(orc, kernel):
    rng = np.random.default_rng(2)
    n = 257
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    sinks = np.array([0, 5, 100, 256])
    a = orc.kick(x, y, z, 1.0 / n, 0.02, kernel, sinks=sinks)
    b = orc.kick(x, y, z, 1.0 / n, 0.02, kernel, sinks=sinks, long_double=True)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-12, atol=1e-14 * np.abs(v).max())


# ------------------------------------------------------------------ epilogue
def _state(n=4):
    return (np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32),
            np.zeros(n, np.int8))


def test_epilogue_half_kick_sign_and_first(orc):
    fx_, fy_, fz_ = (np.array([1.0, -2.0, 0.5, 0.0], np.float32) for _ in range(3))
    vx, vy, vz, rung = _state()
    p = orc.KickParams(ascending=1, descending=0, top=1, first=0, rung=2, t0=1.0, eta=0.1, max_dt=1.0, scale=1.0)
    orc.kick_epilogue(fx_, fy_, fz_, vx, vy, vz, rung, p, 0.01)
    assert np.array_equal(vx, fx_ * np.float32(0.125))  # dt = t0/2 / 2^rung
    vx2, vy2, vz2, rung2 = _state()
    p.top = 0
    orc.kick_epilogue(fx_, fy_, fz_, vx2, vy2, vz2, rung2, p, 0.01)
    assert np.array_equal(vx2, -vx)  # kick.cpp:78: sgn = top ? 1 : -1
    vx3, vy3, vz3, rung3 = _state()
    p.first = 1
    orc.kick_epilogue(fx_, fy_, fz_, vx3, vy3, vz3, rung3, p, 0.01)
    assert not vx3.any()  # kick.cpp:81: no ascending kick on the first step


def test_epilogue_rung_steps_by_at_most_one(orc):
    """kick.cpp:88-93: dt_i = min(eta sqrt(scale h/|F|), t0, max_dt); the rung
    rises by one when t0/dt_i > 2^rung and otherwise falls by at most one."""
    h, eta, t0 = 0.01, 1.0, 1.0
    # |F| chosen so that dt_i = eta sqrt(h/|F|) is exactly t0/8, t0/2, t0
    g = np.array([64 * h, 4 * h, h, 0.0], np.float32)
    fx_ = g.copy()
    fy_ = np.zeros(4, np.float32)
    fz_ = np.zeros(4, np.float32)
    vx, vy, vz, _ = _state()
    rung = np.array([5, 0, 3, 2], np.int8)
    p = orc.KickParams(ascending=0, descending=1, top=1, first=0, rung=1, t0=t0, eta=eta, max_dt=10.0, scale=1.0)
    mr = orc.kick_epilogue(fx_, fy_, fz_, vx, vy, vz, rung, p, h)
    # t0/dt = 8 > 2 -> rung 2 vs old-1=4 -> 4 ; 2 > 2 false -> max(1, -1) = 1 ;
    # 1 > 2 false -> max(1, 2) = 2 ; F = 0: dt = t0 -> max(1, 1) = 1
    assert rung.tolist() == [4, 1, 2, 1]
    assert mr == 4
    assert np.array_equal(vx, g * np.float32(0.25))


def test_epilogue_max_dt_clips(orc):
    fx_ = np.array([1e-8], np.float32)
    vx, vy, vz, rung = _state(1)
    p = orc.KickParams(ascending=0, descending=1, top=1, first=0, rung=0, t0=1.0, eta=1.0, max_dt=0.2, scale=1.0)
    orc.kick_epilogue(fx_, fx_ * 0, fx_ * 0, vx, vy, vz, rung, p, 1.0)
    assert rung[0] == 1  # t0/max_dt = 5 > 2^0


def test_epilogue_rung_test_is_exact_at_the_power_of_two(orc):
    """kick.cpp:92 promotes when ceilf(log2f(t0/dt)) > rung.  Here the test is
    the exact t0/dt > 2^rung.  They differ only where t0/dt lies within a few
    FP32 ulp above 2^rung: there log2f rounds down to rung itself and ct2 does
    not promote.  This case records that intended divergence: t0/dt =
    16 (1 + 2^-23) with rung 4 promotes here, and the FP32 log2 of the same
    quotient is exactly 4.0."""
    t0 = np.float32(1.0)
    q = np.float32(16.0) * (np.float32(1.0) + np.float32(2.0 ** -23))
    dt = np.float32(t0 / q)
    if np.float32(t0 / dt) != q:  # nudge dt by one ulp until the FP32 quotient is q
        dt = np.nextafter(dt, np.float32(0.0)) if np.float32(t0 / dt) < q else np.nextafter(dt, np.float32(1.0))
    assert np.float32(t0 / dt) == q and q > np.float32(16.0)
    assert np.ceil(np.log2(q)) == np.float32(4.0)  # the reference's float evaluation: no promotion
    vx, vy, vz, rung = _state(1)
    rung[:] = 0
    zero = np.zeros(1, np.float32)  # F = 0: dt_i = min(inf, t0, max_dt) = max_dt
    p = orc.KickParams(ascending=0, descending=1, top=1, first=0, rung=4, t0=float(t0), eta=1.0,
                       max_dt=float(dt), scale=1.0)
    orc.kick_epilogue(zero, zero, zero, vx, vy, vz, rung, p, 1.0)
    assert rung[0] == 5  # exact test: 16 (1 + 2^-23) > 2^4
    rung[:] = 0
    p.max_dt = float(np.float32(1.0 / 16.0))  # t0/dt = 16 exactly: no promotion in either form
    orc.kick_epilogue(zero, zero, zero, vx, vy, vz, rung, p, 1.0)
    assert rung[0] == 4


# ------------------------------------------------------------------ drift
def exact_fixed32(s: np.float32) -> int:
    s = float(s)
    if not np.isfinite(s):
        return 0
    return round(Fraction(s) * TWO32) % TWO32  # round() on Fraction: ties to even


def test_fixed32_conversion_is_exact_rne(orc):
    rng = np.random.default_rng(9)
    vals = [0.0, -0.0, 2.0 ** -33, 3 * 2.0 ** -33, -(2.0 ** -33), 0.25, -0.25, 0.75, -0.75, 1.25,
            0.5, -0.5, 2.0 ** 31, 3e9, -3e9, 1e30, 1e-12, float("nan"), float("inf"), 5e-10]
    vals += list(rng.standard_normal(2000) * 1e-3) + list(rng.standard_normal(500) * 10)
    vals += list((rng.integers(-2 ** 40, 2 ** 40, 500) + 0.5) * 2.0 ** -32)  # exact ties
    for v in vals:
        s = np.float32(v)
        assert orc.fixed32_from_float(s) == exact_fixed32(s), v


def test_drift_closed_form_wrap_and_reversibility(orc):
    x = np.array([0, 2 ** 32 - 1, 12345, 2 ** 31], np.uint32)
    y = x.copy()
    z = x.copy()
    v = np.array([0.25, 2.0 ** -32, -0.25, 1.5], np.float32)
    nx, ny, nz = orc.drift(x, y, z, v, v, v, 1.0, 1.0)
    assert nx.tolist() == [2 ** 30, 0, (12345 - 2 ** 30) % 2 ** 32, (2 ** 31 + 2 ** 31) % 2 ** 32]
    rng = np.random.default_rng(4)
    n = 1000
    x = rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32)
    vx = (rng.standard_normal(n) * 0.3).astype(np.float32)
    a = orc.drift(x, x, x, vx, vx, vx, 0.37, 1.7)
    b = orc.drift(*a, vx, vx, vx, -0.37, 1.7)
    assert np.array_equal(b[0], x)  # time reversibility under RNE


def test_drift_matches_exact_model(orc):
    rng = np.random.default_rng(8)
    n = 500
    x = rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32)
    vx = (rng.standard_normal(n)).astype(np.float32)
    dt, scale = np.float32(0.01), np.float32(0.7)
    lam = np.float32(dt / scale)
    nx, _, _ = orc.drift(x, x, x, vx, vx, vx, dt, scale)
    want = [(int(a) + exact_fixed32(np.float32(v * lam))) % TWO32 for a, v in zip(x, vx)]
    assert nx.tolist() == want
