"""The multipole-mesh restatement (particle.cpp:116-271, kick.cpp:11-29)
pinned without sfmm: its Green's-derivative table against the Ewald
oracle's potential and force (themselves pinned by the Madelung constant),
higher derivatives against finite differences of lower ones, the FFT table
against point values, order 1 against a nearest-grid-point mesh built from
the Ewald oracle, and the full force converging to the periodic direct sum
as the order rises."""
import numpy as np
import pytest

TWO32 = 2 ** 32


def to_fixed(p):
    return np.round((np.asarray(p) % 1.0) * TWO32).astype(np.uint64).astype(np.uint32)


def test_component_order(orc):
    idx = orc.mm_index(3)
    assert idx.shape == (20, 3)
    assert idx[:4].tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    assert (idx.sum(1)[:-1] <= idx.sum(1)[1:]).all()
    assert len({tuple(r) for r in idx}) == 20


def pair_ewald(orc, rho, nc):
    """Potential and force (per unit GM, box units) on a sink at 0 from a unit
    source at -rho cells, by the Ewald oracle: two particles minus one."""
    src = to_fixed(-np.asarray(rho) / nc)
    x = np.array([0, src[0]], np.uint32)
    y = np.array([0, src[1]], np.uint32)
    z = np.array([0, src[2]], np.uint32)
    # softening 1e-5: the sink's self term 1/eps cancels in the difference to
    # ~1e-11, and at r > 1/8 the pair's own softened image differs from 1/r by
    # (eps/r)^2 / 2 < 4e-9 (relative)
    two = orc.ewald_kick(x, y, z, 1.0, 1e-5, sinks=[0], alpha=2.5, nimg=3, kmax=8)
    one = orc.ewald_kick(x[:1], y[:1], z[:1], 1.0, 1e-5, sinks=[0], alpha=2.5, nimg=3, kmax=8)
    pot = two[3][0] - one[3][0]
    f = np.array([two[q][0] - one[q][0] for q in range(3)])
    # the exact source position after rounding to fixed point, in cells (minimum image)
    d = np.array([(np.int64(v) + 2 ** 31) % 2 ** 32 - 2 ** 31 for v in src], np.float64) / TWO32 * nc
    return pot, f, -d


@pytest.mark.parametrize("near", [False, True])
def test_green_point_matches_ewald(orc, near):
    nc = 8
    rng = np.random.default_rng(3)
    for _ in range(6):
        rho = rng.uniform(-3.5, 3.5, 3)
        if np.linalg.norm(rho) < 1.0:
            continue
        pot, f, rho = pair_ewald(orc, rho, nc)
        D = orc.mm_green_point(2, nc, rho, near=near)
        r = np.linalg.norm(rho)
        g = D[0] + (1.0 / r if near else 0.0)
        grad = D[1:4] + (-rho / r ** 3 if near else 0.0)
        assert g == pytest.approx(pot / nc, rel=1e-8, abs=1e-11)
        assert np.allclose(grad, f / nc ** 2, rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize("near", [False, True])
def test_green_point_derivatives_by_finite_differences(orc, near):
    """d D_c / d rho_q = D_{c + e_q} for every component below the top order."""
    P, nc, eps = 6, 8, 1e-4
    idx = orc.mm_index(P)
    lookup = {tuple(e): c for c, e in enumerate(idx)}
    rng = np.random.default_rng(11)
    for _ in range(3):
        rho = rng.uniform(-3.0, 3.0, 3)
        if not near and np.linalg.norm(rho) < 1.5:
            rho *= 2.0
        D = orc.mm_green_point(P, nc, rho, near=near)
        for q in range(3):
            e = np.zeros(3)
            e[q] = eps
            Dp = orc.mm_green_point(P, nc, rho + e, near=near)
            Dm = orc.mm_green_point(P, nc, rho - e, near=near)
            fd = (Dp - Dm) / (2 * eps)
            for c, ex in enumerate(idx):
                if ex.sum() >= P:
                    continue
                up = list(ex)
                up[q] += 1
                want = D[lookup[tuple(up)]]
                scale = np.abs(D[idx.sum(1) == ex.sum() + 1]).max()
                assert abs(fd[c] - want) <= 2e-6 * scale, (rho, q, ex, fd[c], want)


def test_green_table_matches_point_values(orc):
    """The inverse transform of the table (times nc^3, with the parity phase)
    gives D_c at every cell offset, near offsets without their nearest image."""
    P, nc = 4, 8
    idx = orc.mm_index(P)
    Dt = orc.mm_green_table(P, nc)
    deg = idx.sum(1)
    samples = [(0, 0, 0), (1, 0, 0), (3, 1, 0), (2, 2, 2), (-4, 3, -1), (4, 4, 4), (-1, -3, 2)]
    back = [np.real(np.fft.ifftn(Dt[c] * nc ** 3 * (1j if deg[c] % 2 else 1.0))) for c in range(len(idx))]
    for s in samples:
        rho = np.array(s, float)
        w = [((v + nc // 2) % nc) - nc // 2 for v in s]
        near = w[0] ** 2 + w[1] ** 2 + w[2] ** 2 <= 10
        D = orc.mm_green_point(P, nc, np.array(w, float), near=near)
        got = np.array([back[c][s[0] % nc, s[1] % nc, s[2] % nc] for c in range(len(idx))])
        assert np.allclose(got, D, rtol=1e-9, atol=1e-11 * np.abs(D).max()), (s, np.abs(got - D).max())
        # parity: D_c(-rho) = (-1)^deg D_c(rho)
        neg = np.array([back[c][-s[0] % nc, -s[1] % nc, -s[2] % nc] for c in range(len(idx))])
        assert np.allclose(neg, got * (-1.0) ** deg, rtol=1e-9, atol=1e-11 * np.abs(D).max())


def random_set(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 3))
    if clustered:
        c = rng.random((4, 3))
        k = n // 2
        p[:k] = c[rng.integers(0, 4, k)] + 0.03 * rng.standard_normal((k, 3))
    return [to_fixed(p[:, q]) for q in range(3)]


def test_order_one_is_a_nearest_grid_point_mesh(orc):
    """At P = 1 the far force is the sum over cells of (cell mass) x grad G_far
    at the offset of the cell centres: the Ewald force between cell centres
    (one Ewald oracle call per cell offset) less, for the near set, the
    Newtonian force; a cell's own particles add nothing (zero by symmetry)."""
    nc, n = 8, 300
    x, y, z = random_set(n, 5)
    fx, fy, fz, _ = orc.mm_forces(x, y, z, nc, 1, 1.0, 1e-12, parts=2)
    sh = 32 - 3
    cells = np.stack([x >> sh, y >> sh, z >> sh], 1).astype(np.int64)
    half = np.uint32(1 << (sh - 1))
    g = {}
    for d in np.ndindex(nc, nc, nc):
        if d == (0, 0, 0):
            continue
        px = np.array([half, (d[0] << sh) + half], np.uint32)
        py = np.array([half, (d[1] << sh) + half], np.uint32)
        pz = np.array([half, (d[2] << sh) + half], np.uint32)
        e = orc.ewald_kick(px, py, pz, 1.0, 1e-12, sinks=[1], alpha=2.5, nimg=3, kmax=8)
        w = np.array([((v + nc // 2) % nc) - nc // 2 for v in d], float)
        f = np.array([e[0][0], e[1][0], e[2][0]])
        if (w ** 2).sum() <= 10:  # the nearest image is summed directly
            r = np.linalg.norm(w) / nc
            f -= (-w / nc) / r ** 3
        g[d] = f
    want = np.zeros((3, n))
    for i in range(n):
        for j in range(n):
            d = tuple(int(v) for v in (cells[i] - cells[j]) % nc)
            if d != (0, 0, 0):
                want[:, i] += g[d]
    got = np.stack([fx, fy, fz])
    assert np.allclose(got, want, rtol=1e-8, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("clustered", [False, True])
def test_force_converges_to_the_periodic_direct_sum(orc, clustered):
    """Near-set pairs plus the far field against the Ewald direct sum: the
    median force error falls about tenfold with every added pair of orders
    (uniform: 1.4e-2, 1.2e-3, 9.6e-5, 1.0e-5 at P = 2, 4, 6, 8).  The
    potential levels off near 1e-6: the Ewald oracle softens every pair's
    nearest image, the mesh only the direct pairs (by (eps/r)^2 ~ 1e-5 at
    r ~ 0.3)."""
    nc, n = 8, 400
    x, y, z = random_set(n, 7, clustered)
    gm, soft = 1.0 / n, 1e-3
    ex, ey, ez, ep = orc.ewald_kick(x, y, z, gm, soft, alpha=2.5, nimg=3, kmax=8)
    J = np.stack([ex, ey, ez])
    med, pmed = [], []
    for P in (2, 4, 6, 8):
        fx, fy, fz, pot = orc.mm_forces(x, y, z, nc, P, gm, soft)
        e = np.linalg.norm(np.stack([fx, fy, fz]) - J, axis=0) / np.linalg.norm(J, axis=0)
        med.append(np.median(e))
        pmed.append(np.median(np.abs(pot - ep) / np.abs(ep)))
    print("median force error by order 2,4,6,8:", med, "potential:", pmed)
    assert all(b < 0.5 * a for a, b in zip(med, med[1:])), med
    assert med[-1] < (5e-6 if clustered else 2e-5)
    assert pmed[-1] < 3e-6


def test_near_part_is_the_direct_sum_over_the_near_set(orc):
    """parts=1 is the softened minimum-image sum over the cells within
    |offset|^2 <= 10; it includes the self pair's potential and, with the
    far field, reproduces the full decomposition."""
    nc, n = 8, 200
    x, y, z = random_set(n, 9)
    gm, soft = 1.0 / n, 2e-3
    a = orc.mm_forces(x, y, z, nc, 4, gm, soft, parts=1)
    b = orc.mm_forces(x, y, z, nc, 4, gm, soft, parts=2)
    c = orc.mm_forces(x, y, z, nc, 4, gm, soft, parts=3)
    for q in range(4):
        assert np.allclose(a[q] + b[q], c[q], rtol=1e-12, atol=1e-14)
    sh = 29
    cells = np.stack([x >> sh, y >> sh, z >> sh], 1).astype(np.int64)
    for i in (0, 17, 123):
        dc = (cells - cells[i] + nc // 2) % nc - nc // 2
        far = (dc ** 2).sum(1) > 10
        # the direct sum over the far cells only, from the all-pairs oracle with those sources
        sel = np.flatnonzero(~far)
        fx, fy, fz, p = orc.kick(x[sel], y[sel], z[sel], gm, soft, orc.KERNEL_PLUMMER,
                                 sinks=[int(np.flatnonzero(sel == i)[0])])
        assert np.allclose([a[0][i], a[1][i], a[2][i], a[3][i]], [fx[0], fy[0], fz[0], p[0]], rtol=1e-12)
