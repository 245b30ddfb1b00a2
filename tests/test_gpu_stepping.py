"""GPU checks of the stepping configuration's pieces: the Gaussian-random-field
initial conditions (against an FP64 numpy construction from the same noise
model's statistics) and the leapfrog loop (linear growth of a single mode
against the Einstein-de Sitter solution)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2 ** 32


def displacement(p, side):
    idx = np.arange(side ** 3)
    out = []
    for a, v in enumerate((p.x, p.y, p.z)):
        q = ((idx // side ** (2 - a)) % side + 0.5) / side
        d = ((v.astype(np.int64) - np.round(q * TWO32).astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31) / TWO32
        out.append(d)
    return np.stack(out)


def test_grf_ics_statistics(ct2):
    side, rms = 64, 0.3
    ct2.particle_allocate(side ** 3)
    ct2.ic_grf(side, rms, -2.0, 5, 2.0)
    p = ct2.particles_download()
    d = displacement(p, side) * side  # in cells
    assert np.sqrt(np.mean(d ** 2)) == pytest.approx(rms, rel=1e-6)
    assert np.allclose(p.vx, 2.0 * d[0] / side, rtol=1e-5, atol=1e-9)
    # curl-free: the displacement spectrum is longitudinal, psi_hat || k
    F = [np.fft.fftn(d[a].reshape(side, side, side)) for a in range(3)]
    m = np.fft.fftfreq(side, 1 / side)
    MX, MY, MZ = np.meshgrid(m, m, m, indexing="ij")
    cross = np.abs(MY * F[2] - MZ * F[1]) + np.abs(MZ * F[0] - MX * F[2])
    assert cross.max() < 1e-4 * np.abs(F[0]).max() * side
    # density P(k) ~ k^-2 means psi power ~ k^-4: compare two shells
    P = np.abs(F[0]) ** 2 + np.abs(F[1]) ** 2 + np.abs(F[2]) ** 2
    kk = np.sqrt(MX ** 2 + MY ** 2 + MZ ** 2)
    p4 = P[(kk > 3.5) & (kk < 4.5)].mean()
    p8 = P[(kk > 7.5) & (kk < 8.5)].mean()
    assert p4 / p8 == pytest.approx(16.0, rel=0.15)
    # cut at the Nyquist sphere
    assert P[kk >= side / 2].max() < 1e-12 * P.max()
    # seeded
    ct2.ic_grf(side, rms, -2.0, 5, 2.0)
    q = ct2.particles_download()
    assert np.array_equal(p.x, q.x)


def test_leapfrog_linear_growth(ct2):
    """A single long-wavelength mode in an Einstein-de Sitter universe.  The PM
    force of the mode is f 4 pi psi with f = exp(-k^2 rs^2); linear theory then
    gives D = c+ (a/a0)^p+ + c- (a/a0)^p-, p = (-1/2 +- sqrt(1/4 + 6 f))/2, with
    c+- fixed by the growing-mode start D(a0) = 1, a0 D'(a0) = 1.  The KDK loop
    with FP64 factors must reproduce D(0.03)."""
    side = 32
    n = side ** 3
    H = np.sqrt(8 * np.pi / 3)
    A0 = 1e-4
    q = (np.arange(side) + 0.5) / side
    X, Y, Z = np.meshgrid(q, q, q, indexing="ij")
    psi = A0 * np.sin(2 * np.pi * X.ravel())
    ct2.particle_allocate(n)
    ct2.particles_upload((((X.ravel() + psi) % 1) * TWO32).astype(np.uint64).astype(np.uint32),
                         (Y.ravel() * TWO32).astype(np.uint64).astype(np.uint32),
                         (Z.ravel() * TWO32).astype(np.uint64).astype(np.uint32),
                         (H * 0.02 ** 0.5 * psi).astype(np.float32), np.zeros(n, np.float32),
                         np.zeros(n, np.float32), np.zeros(n, np.int8))
    ct2.set_force_options(1.0 / (50 * side), 1.0 / n, ct2.KERNEL_PLUMMER)
    rs = 1.0
    ct2.pm_init(side, rs)
    cosmo = ct2.cosmology(H)
    a = np.linspace(0.02, 0.03, 21)
    ah = 0.5 * (a[:-1] + a[1:])
    pending = ct2.cosmos_kick_factor(cosmo, a[0], ah[0])
    for i in range(20):
        ct2.kick_pm(ct2.kick_params(ascending=True, top=True, rung=0, t0=float(2 * pending)))
        ct2.drift(float(ct2.cosmos_drift_factor(cosmo, a[i], a[i + 1])), 1.0)
        pending = ct2.cosmos_kick_factor(cosmo, ah[i], a[i + 1])
        if i + 1 < 20:
            pending += ct2.cosmos_kick_factor(cosmo, a[i + 1], ah[i + 1])
    p = ct2.particles_download()
    d = (((p.x.astype(np.int64) - (X.ravel() * TWO32).astype(np.int64) + 2 ** 31) % 2 ** 32) - 2 ** 31) / TWO32
    f = np.exp(-(2 * np.pi * rs / side) ** 2)
    pp, pm = (-0.5 + np.sqrt(0.25 + 6 * f)) / 2, (-0.5 - np.sqrt(0.25 + 6 * f)) / 2
    cp = (1 - pm) / (pp - pm)
    growth = cp * 1.5 ** pp + (1 - cp) * 1.5 ** pm
    k = np.argmax(psi)
    assert d[k] / psi[k] == pytest.approx(growth, rel=2e-3)
    assert abs(growth - 1.5) > 0.004  # the filter's 0.5 % effect is resolved by the check


@pytest.mark.parametrize("method", ["pm_mlp", "pm_near", "mm"])
def test_config5_steps(ct2, orc, method):
    """The stepping configuration at full size: 256^3 particles from a
    P(k) ~ k^-2 Gaussian random field (rms 0.3 cell), a 512^3 TSC grid, KDK
    steps from a = 0.02 with da = 0.001 (the 10-step schedule to 0.03; the
    first 3 steps here), FP64 kick and drift factors.  Each step the force is
    judged on 16,384 sampled particles against the periodic FP64 direct sum:
    PM + the committed learned correction (mlp_grf_ng2x.npz, its own mesh)
    at median <= 3e-3, PM + near field (rs 2 cells, cutoff 5.5 rs) at
    median <= 1e-3, ct2's multipole mesh (128^3 cells, order 6) at median
    <= 3e-4; each drift is bit-exact against the oracle's on a sample."""
    import os

    from ct2_amd.train_mlp import load_weights

    side, ng, steps = 256, 512, 3
    n = side ** 3
    H = float(np.sqrt(8 * np.pi / 3))
    cosmo = ct2.cosmology(H)
    ct2.particle_allocate(n)
    ct2.ic_grf(side, 0.3, -2.0, 1, H * 0.02 ** 0.5)
    ct2.set_force_options(float(np.float32(1.0 / (50 * side))), float(np.float32(1.0 / n)), ct2.KERNEL_PLUMMER)
    if method == "pm_mlp":
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        w, meta = load_weights(os.path.join(root, "ct2_amd", "weights", "mlp_grf_ng2x.npz"))
        assert meta["ng"] == 2 * meta["side"]  # trained on a grid of twice the lattice, as here
        ct2.pm_assignment(int(meta["order"]))
        ct2.pm_init(ng, float(meta["rs_cells"]))
        ct2.pm_near(0.0)
        ct2.mlp_init(1)
        ct2.mlp_set_weights(w)
        ct2.mlp_set_mesh(*ct2.mlp_mesh_of(meta))
        kick, limit = ct2.kick_pm_mlp, 3e-3
    elif method == "mm":
        ct2.mm_init(ct2.mm_cells_for(n), 6)
        kick, limit = ct2.kick_mm, 3e-4
    else:
        ct2.pm_assignment(2)
        ct2.pm_init(ng, 2.0)
        ct2.pm_near(5.5)
        kick, limit = ct2.kick_pm, 1e-3
    try:
        a = 0.02 + 0.001 * np.arange(steps + 1)
        ah = 0.5 * (a[:-1] + a[1:])
        rng = np.random.default_rng(7)
        pending = ct2.cosmos_kick_factor(cosmo, a[0], ah[0])
        for i in range(steps):
            kick(ct2.kick_params(ascending=True, top=True, rung=0, t0=float(2.0 * pending)), ct2.KICK_STORE)
            assert ct2.last_force_path() == {"pm_mlp": ct2.PATH_PM_MLP, "pm_near": ct2.PATH_PM_NEAR,
                                             "mm": ct2.PATH_MM}[method]
            idx = np.sort(rng.choice(n, 16384, replace=False))
            fx, fy, fz, _ = ct2.forces_download()
            jx, jy, jz, _ = ct2.kick_sample_fp64(idx, periodic=True)
            J = np.stack([jx, jy, jz])
            e = np.linalg.norm(np.stack([fx[idx], fy[idx], fz[idx]]) - J, axis=0) / np.linalg.norm(J, axis=0)
            print(f"config 5 {method} step {i}: median {np.median(e):.3e} p90 {np.percentile(e, 90):.3e}",
                  flush=True)
            assert np.median(e) <= limit, (i, np.median(e))
            before = ct2.particles_download()
            D = float(ct2.cosmos_drift_factor(cosmo, a[i], a[i + 1]))
            ct2.drift(D, 1.0)
            after = ct2.particles_download()
            s = idx[:4096]
            wx, wy, wz = orc.drift(before.x[s], before.y[s], before.z[s], before.vx[s], before.vy[s], before.vz[s],
                                   D, 1.0)
            assert np.array_equal(after.x[s], wx) and np.array_equal(after.y[s], wy) and np.array_equal(after.z[s], wz)
            if i + 1 < steps:
                pending = (ct2.cosmos_kick_factor(cosmo, ah[i], a[i + 1])
                           + ct2.cosmos_kick_factor(cosmo, a[i + 1], ah[i + 1]))
    finally:
        ct2.pm_near(0.0)
        ct2.pm_assignment(1)
