"""The oracle's near-field complement of the PM force (kick.cpp:105-125
structure: direct pairs over the neighbourhood, mesh beyond).  Pinned by the
periodic FP64 direct sum it must converge to, by its own limits and by
cutoff independence, not by restating its formula."""
import numpy as np
import pytest

TWO32 = 2 ** 32


def lattice(side, sigma, seed):
    rng = np.random.default_rng(seed)
    q = (np.arange(side) + 0.5) / side
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()]) + rng.standard_normal((3, side ** 3)) * sigma / side
    return [np.round((p % 1.0) * TWO32).astype(np.uint64).astype(np.uint32) for p in P]


def rel(F, J):
    return np.linalg.norm(F - J, axis=0) / np.linalg.norm(J, axis=0)


@pytest.fixture(scope="module")
def case(orc):
    side, ng = 16, 32
    x, y, z = lattice(side, 0.2, 1)
    n = x.size
    gm, soft = 1.0 / n, 1.0 / (50 * side)
    sinks = np.random.default_rng(3).choice(n, 48, replace=False)
    J = np.stack(orc.ewald_kick(x, y, z, gm, soft, sinks=sinks))
    return dict(side=side, ng=ng, x=x, y=y, z=z, gm=gm, soft=soft, sinks=sinks, J=J)


def p3m(orc, c, rs_cells, rcut_rs, kernel=0):
    rs = rs_cells / c["ng"]
    pm = orc.pm_forces(c["x"], c["y"], c["z"], c["ng"].bit_length() - 1, c["gm"], rs_cells)
    nr = orc.pm_near(c["x"], c["y"], c["z"], c["gm"], c["soft"], rs, rcut_rs * rs, kernel, sinks=c["sinks"])
    return np.stack([pm[a][c["sinks"]] + nr[a] for a in range(4)]), np.stack([pm[a][c["sinks"]] for a in range(4)])


def test_pm_plus_near_field_converges_to_periodic_sum(orc, case):
    """PM alone misses the force almost entirely on a near-uniform lattice
    (the near neighbours dominate); PM + near field reaches the periodic FP64
    direct sum to the mesh's discretisation error, which falls as the split
    radius grows in mesh cells."""
    J = case["J"]
    errs = []
    for rs_cells in (1.25, 2.0, 3.0):
        F, pm = p3m(orc, case, rs_cells, 5.0)
        e = rel(F[:3], J[:3])
        errs.append(np.median(e))
        assert np.median(rel(pm[:3], J[:3])) > 0.5
    assert errs[0] < 3e-2 and errs[2] < 3e-3, errs
    assert errs[0] > errs[1] > errs[2]


def test_potential_matches_periodic_convention(orc, case):
    """The background term makes PM + near-field potential the periodic
    (zero-mean) potential of the Ewald sum; without it the two would differ
    by GM N 4 pi rs^2 (0.11 here, 60 % of the potential); what remains is
    the mesh's discretisation error, < 1e-3 of the potential."""
    F, _ = p3m(orc, case, 3.0, 5.0)
    J = case["J"]
    assert np.max(np.abs(F[3] - J[3])) < 1e-3 * np.abs(J[3]).max()


def test_cutoff_independence(orc, case):
    """Beyond ~5 rs the filtered and unfiltered kernels agree to erfc-level
    terms: moving the cutoff from 5.5 rs to 7 rs changes the force by
    ~1e-4 of its size."""
    a, _ = p3m(orc, case, 2.0, 5.5)
    b, _ = p3m(orc, case, 2.0, 7.0)
    assert np.median(rel(a[:3], b[:3])) < 5e-4


def test_two_body_and_self_limits(orc):
    """Two particles closer than the cutoff: the near field restores the
    softened pair force exactly where the mesh kernel is negligible
    (r << rs) and vanishes beyond the cutoff; the self pair adds
    1/eps - 1/(rs sqrt(pi)) to the potential and no force."""
    gm, soft, rs = 1.0, 1e-3, 0.05
    base = np.uint32(1 << 31)
    for r in (1e-4, 3e-3, 0.4 * rs):
        d = np.uint32(round(r * TWO32))
        x = np.array([base, base + d], np.uint32)
        y = np.array([base, base], np.uint32)
        ax, ay, az, pot = orc.pm_near(x, y, y.copy(), gm, soft, rs, 6 * rs)
        rr = int(d) / TWO32
        plummer = rr * (rr * rr + soft * soft) ** -1.5
        # what is removed is the mesh's pair force: for r << rs that of a mass
        # spread as the Gaussian (4 pi rs^2)^-3/2 exp(-r^2 / 4 rs^2), i.e.
        # (4 pi / 3) rho(0) r = r / (6 sqrt(pi) rs^3)
        assert (plummer - ax[0]) / rr == pytest.approx(1 / (6 * np.sqrt(np.pi) * rs ** 3),
                                                       rel=0.25 * (rr / rs) ** 2 + 1e-9)
        assert ax[1] == -ax[0] and ay[0] == 0.0
    x = np.array([base], np.uint32)
    ax, ay, az, pot = orc.pm_near(x, x, x, gm, soft, rs, 6 * rs)
    assert ax[0] == 0.0
    assert pot[0] == pytest.approx(1 / soft - 1 / (rs * np.sqrt(np.pi)) - 4 * np.pi * rs * rs, rel=1e-14)
    # beyond the cutoff: only the self terms remain
    d = np.uint32(round(0.31 * TWO32))
    x = np.array([base, base + d], np.uint32)
    y = np.array([base, base], np.uint32)
    ax, ay, az, pot = orc.pm_near(x, y, y.copy(), gm, soft, rs, 6 * rs)
    assert ax[0] == 0.0
    assert pot[0] == pytest.approx(1 / soft - 1 / (rs * np.sqrt(np.pi)) - 8 * np.pi * rs * rs, rel=1e-14)


def test_series_branch_is_continuous(orc):
    """The small-separation series and the closed form meet at u = 1e-3."""
    gm, soft, rs = 1.0, 1e-6, 0.01
    base = np.uint32(1 << 31)
    vals = []
    for r in (2 * rs * 0.999e-3, 2 * rs * 1.001e-3):
        d = np.uint32(round(r * TWO32))
        x = np.array([base, base + d], np.uint32)
        y = np.array([base, base], np.uint32)
        ax, _, _, pot = orc.pm_near(x, y, y.copy(), gm, soft, rs, 6 * rs)
        rr = int(d) / TWO32
        plummer = rr * (rr * rr + soft * soft) ** -1.5
        vals.append((plummer - ax[0]) / rr)  # the filtered kernel divided by r
    assert vals[0] == pytest.approx(vals[1], rel=1e-3)
    assert vals[0] == pytest.approx(2 / np.sqrt(np.pi) * (2 / 3) / (2 * rs) ** 3, rel=1e-5)


def test_tsc_mesh_is_ten_times_more_accurate(orc, case):
    """With the quadratic-spline assignment the mesh's discretisation error at
    a given split radius falls by about an order of magnitude."""
    J = case["J"]
    rs_cells = 2.0
    rs = rs_cells / case["ng"]
    nr = orc.pm_near(case["x"], case["y"], case["z"], case["gm"], case["soft"], rs, 5.5 * rs, sinks=case["sinks"])
    errs = []
    for order in (1, 2):
        pm = orc.pm_forces(case["x"], case["y"], case["z"], case["ng"].bit_length() - 1, case["gm"], rs_cells,
                           order=order)
        F = np.stack([pm[a][case["sinks"]] + nr[a] for a in range(3)])
        errs.append(np.median(rel(F, J[:3])))
    assert errs[1] < 1e-3 and errs[1] < errs[0] / 5, errs
