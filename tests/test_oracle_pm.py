"""Pins the FP64 particle-mesh and Ewald oracles (particle.cpp:116-271 shape;
the periodic force of particle.cpp:141-144) with closed forms, invariants and
each other."""
import numpy as np
import pytest

TWO32 = 2 ** 32
XI_SC = -2.837297479480620  # Madelung/Wigner constant of the simple cubic lattice with background


def fx(v):
    return (np.asarray(v, dtype=float) % 1.0 * TWO32).astype(np.uint64).astype(np.uint32)


# ------------------------------------------------------------------ CIC
def test_cic_conserves_mass_and_places_nodes(orc):
    rng = np.random.default_rng(3)
    n = 1000
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    W = orc.cic_deposit(x, y, z, 4)
    # each particle deposits 2^32 minus at most 8 units of truncation
    assert n * TWO32 - 8 * n <= int(W.sum()) <= n * TWO32
    # a particle on a node puts all of its weight there
    W1 = orc.cic_deposit(fx([3 / 16]), fx([5 / 16]), fx([0.0]), 4)
    assert W1[3, 5, 0] == TWO32 and W1.sum() == TWO32
    # halfway between two nodes along x: equal halves
    W2 = orc.cic_deposit(fx([3.5 / 16]), fx([5 / 16]), fx([0.0]), 4)
    assert W2[3, 5, 0] == W2[4, 5, 0] == TWO32 // 2
    # periodic wrap: the upper neighbour of the last node is node 0
    W3 = orc.cic_deposit(fx([15.25 / 16]), fx([0.0]), fx([0.0]), 4)
    assert W3[15, 0, 0] == 3 * TWO32 // 4 and W3[0, 0, 0] == TWO32 // 4


def test_green_table_closed_form(orc):
    ng, gm = 32, 0.5
    t = orc.pm_green_table(ng, gm, 0.0)
    assert t[0, 0, 0] == 0.0
    s = np.sin(np.pi / ng) / (np.pi / ng)
    assert t[1, 0, 0] == pytest.approx(-4 * np.pi * gm / ((2 * np.pi) ** 2 * s ** 4), rel=1e-14)
    assert t[0, 0, 1] == t[1, 0, 0] == t[0, 1, 0]
    tg = orc.pm_green_table(ng, gm, 1.25)
    k2 = 2 * (2 * np.pi) ** 2
    s2 = s ** 2
    want = -4 * np.pi * gm * np.exp(-k2 * (1.25 / ng) ** 2) / (k2 * s2 ** 4)
    assert tg[1, 1, 0] == pytest.approx(want, rel=1e-14)


# ------------------------------------------------------------------ Ewald
@pytest.mark.parametrize("alpha,nimg,kmax", [(2.5, 2, 5), (3.5, 2, 6), (1.8, 3, 4)])
def test_ewald_madelung_constant(orc, alpha, nimg, kmax):
    """A single particle's periodic self-potential is the simple-cubic
    Madelung constant (plus the softened self term 1/eps)."""
    x = np.array([123456789], np.uint32)
    p = orc.ewald_kick(x, x, x, 1.0, 1e3, alpha=alpha, nimg=nimg, kmax=kmax)[3][0]
    assert p - 1e-3 == pytest.approx(XI_SC, abs=1e-12)


def test_ewald_lattice_force_and_potential(orc):
    n = 4
    g = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    soft = 1e-6
    ax, ay, az, p = orc.ewald_kick(fx(X.ravel()), fx(Y.ravel()), fx(Z.ravel()), 1 / n ** 3, soft, sinks=[0, 21, 63])
    assert np.max(np.abs([ax, ay, az])) < 1e-9
    assert np.allclose(p - 1 / (n ** 3 * soft), XI_SC / n ** 2, atol=1e-9)


def test_ewald_symmetry_points_and_alpha_independence(orc):
    y = np.zeros(2, np.uint32)
    # half a box apart along x the periodic pull vanishes
    f = orc.ewald_kick(np.array([0, 2 ** 31], np.uint32), y, y, 1.0, 1e-7)[0]
    assert np.max(np.abs(f)) < 1e-9
    rng = np.random.default_rng(6)
    n = 60
    x, yy, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    a = orc.ewald_kick(x, yy, z, 1 / n, 1e-3, sinks=[0, 7, 33], alpha=2.5, nimg=2, kmax=5)
    b = orc.ewald_kick(x, yy, z, 1 / n, 1e-3, sinks=[0, 7, 33], alpha=3.5, nimg=2, kmax=6)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-10, atol=1e-10)


def test_ewald_close_pair_is_newtonian(orc):
    """At separations far below the box the periodic force is the direct one
    plus the smooth background term -(4 pi/3) d."""
    d = 1e-3
    x = fx([0.3, 0.3 + d])
    y = fx([0.5, 0.5])
    a = orc.ewald_kick(x, y, y, 1.0, 1e-5)[0][0]
    dd = (int(x[1]) - int(x[0])) / TWO32
    direct = dd / (dd * dd + 1e-10) ** 1.5
    assert a == pytest.approx(direct - 4 * np.pi / 3 * dd, rel=1e-9)


def test_linear_single_mode_force(orc):
    """Lattice displaced by a small single Fourier mode psi_x = A sin(2 pi q_x):
    in linear theory the periodic force is 4 pi G rho_bar psi (rho_bar = 1);
    the PM keeps the long-range part, 4 pi psi exp(-k^2 rs^2).  Pins the PM
    normalisation (4 pi, GM, FFT scalings, filter) and the Ewald sum."""
    n, A, g, rs = 32, 1e-4, 5, 1.0
    q = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    psi = A * np.sin(2 * np.pi * X.ravel())
    x, y, z = fx(X.ravel() + psi), fx(Y.ravel()), fx(Z.ravel())
    N = n ** 3
    fxp, fyp, fzp, _ = orc.pm_forces(x, y, z, g, 1.0 / N, rs)
    want = 4 * np.pi * psi
    filt = np.exp(-(2 * np.pi * rs / 2 ** g) ** 2)
    assert np.allclose(fxp, want * filt, rtol=0, atol=2e-3 * np.abs(want).max())
    assert np.max(np.abs(fyp)) < 1e-6 * np.abs(want).max()
    sinks = np.array([5, 700, 3001, np.argmax(psi)])
    e = orc.ewald_kick(x, y, z, 1.0 / N, 1e-4, sinks=sinks)
    # discreteness of the particle lattice: O((k a)^2) ~ 1e-3
    assert np.allclose(e[0], want[sinks], rtol=0, atol=3e-3 * np.abs(want).max())


# ------------------------------------------------------------------ PM
def test_pm_pair_matches_ewald_at_large_separation(orc):
    x = np.array([0, 2 ** 30], np.uint32)
    y = np.array([2 ** 20, 2 ** 20], np.uint32)
    e = orc.ewald_kick(x, y, y, 1.0, 1e-6)
    pm = orc.pm_forces(x, y, y, 6, 1.0, 1.25)
    assert pm[0][0] == pytest.approx(e[0][0], rel=1e-3)
    assert pm[0][1] == pytest.approx(e[0][1], rel=1e-3)


def test_pm_conserves_momentum(orc):
    rng = np.random.default_rng(12)
    n = 500
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    f = orc.pm_forces(x, y, z, 5, 1.0 / n, 1.25)
    for a in f[:3]:
        assert abs(a.sum()) < 1e-10 * np.abs(a).sum()


# ------------------------------------------------------------------ TSC
def test_tsc_conserves_mass_and_places_nodes(orc):
    """The quadratic spline: on a node 1/8, 3/4, 1/8 per axis; half a cell off,
    1/2, 1/2, 0; continuous where the nearest node changes; periodic."""
    rng = np.random.default_rng(4)
    n = 1000
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    W = orc.mesh_deposit(x, y, z, 4, 2)
    assert n * TWO32 - 27 * n <= int(W.sum()) <= n * TWO32
    W1 = orc.mesh_deposit(fx([3 / 16]), fx([5 / 16]), fx([0.0]), 4, 2).astype(float) / TWO32
    w = np.array([1 / 8, 3 / 4, 1 / 8])
    want = np.zeros((16, 16, 16))
    for a, i in enumerate((2, 3, 4)):
        for b, j in enumerate((4, 5, 6)):
            for c, k in enumerate((15, 0, 1)):
                want[i, j, k] = w[a] * w[b] * w[c]
    assert np.allclose(W1, want, atol=1e-9)
    # just below and at the half-cell point between nodes 3 and 4 along x
    lo = np.array([np.uint32(round(3.5 / 16 * TWO32) - 1)])
    hi = np.array([np.uint32(round(3.5 / 16 * TWO32))])
    Wl = orc.mesh_deposit(lo, fx([5 / 16]), fx([0.0]), 4, 2).astype(float) / TWO32
    Wh = orc.mesh_deposit(hi, fx([5 / 16]), fx([0.0]), 4, 2).astype(float) / TWO32
    assert np.abs(Wl - Wh).max() < 1e-5
    assert Wh[3].sum() == pytest.approx(0.5, abs=1e-6) and Wh[4].sum() == pytest.approx(0.5, abs=1e-6)
    assert Wh[2].sum() == 0 and Wh[5].sum() == 0


def test_tsc_green_window(orc):
    ng, gm = 32, 0.5
    s = np.sin(np.pi / ng) / (np.pi / ng)
    t = orc.pm_green_table(ng, gm, 0.0, order=2)
    assert t[1, 0, 0] == pytest.approx(-4 * np.pi * gm / ((2 * np.pi) ** 2 * s ** 6), rel=1e-14)


def test_tsc_linear_mode_and_momentum(orc):
    """The single-mode force and momentum conservation hold for TSC as well;
    lattice aliasing leaves 0.2-0.3 % for any lattice offset against the grid
    (CIC: 0.03 % on the half-node offset of the CIC test, 4-8 % on others)."""
    n, A, g, rs = 32, 1e-4, 5, 1.0
    q = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    psi = A * np.sin(2 * np.pi * X.ravel())
    fxp, fyp, _, _ = orc.pm_forces(fx(X.ravel() + psi), fx(Y.ravel()), fx(Z.ravel()), g, 1.0 / n ** 3, rs, order=2)
    want = 4 * np.pi * psi * np.exp(-(2 * np.pi * rs / 2 ** g) ** 2)
    assert np.allclose(fxp, want, rtol=0, atol=5e-3 * np.abs(want).max())
    rng = np.random.default_rng(12)
    x, y, z = (rng.integers(0, TWO32, 500, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    f = orc.pm_forces(x, y, z, 5, 1.0 / 500, 1.25, order=2)
    for a in f[:3]:
        assert abs(a.sum()) < 1e-10 * np.abs(a).sum()
