"""GPU parity of the learned correction's BF16-MFMA kernels against the FP64
oracle (forward, gradients, features) and the gain test: the trained
network must reduce the PM force error against the periodic FP64 judge."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2 ** 32


def features(rng, m):
    X = np.zeros((m, 128), np.float32)
    X[:, :108] = rng.standard_normal((m, 108))
    X[:, 108:111] = rng.uniform(-0.5, 0.5, (m, 3))
    return X


def test_mlp_forward_vs_oracle(ct2, orc):
    ct2.mlp_init(7)
    w = ct2.mlp_get_weights().astype(np.float64)
    rng = np.random.default_rng(1)
    X = features(rng, 384)
    got = ct2.mlp_forward(X)
    want = orc.mlp_forward(w, X, bf16=True)[:, :3]
    # bf16 activations: an FP32 pre-activation that lands next to a bf16
    # rounding boundary may round one ulp (2^-8) apart from the FP64 one
    err = np.abs(got - want)
    scale = np.sqrt(np.mean(want ** 2))
    assert np.median(err) < 2e-3 * scale and np.max(err) < 3e-2 * scale, (np.median(err) / scale, err.max() / scale)
    exact = orc.mlp_forward(w, X, bf16=False)[:, :3]
    assert np.median(np.abs(got - exact)) < 2e-2 * scale  # bf16 network vs FP64 network


@pytest.mark.parametrize("m", [256, 2048, 32768])
def test_mlp_gradients_vs_oracle(ct2, orc, m):
    """m = 2048 takes the split-K weight gradients (8 K slices and the
    ordered reduction of their planes); 256 the single-slice products;
    32,768 the 256 x 128 LDS-staged tiles (hidden-layer products and the
    layer-2 weight gradient, 128 K slices)."""
    ct2.mlp_init(11)
    w = ct2.mlp_get_weights().astype(np.float64)
    rng = np.random.default_rng(2)
    X = features(rng, m)
    T = np.concatenate([rng.standard_normal((m, 3)), rng.uniform(0.5, 2.0, (m, 1))], axis=1).astype(np.float32)
    loss, g = ct2.mlp_grad(X, T)
    oloss, og = orc.mlp_grad(w, X, T, bf16=True)
    assert loss == pytest.approx(oloss, rel=1e-2)
    # per-matrix relative L2 error; bf16 output gradients bound it at ~1e-2
    offs = [0, 32768, 33024, 98560, 98816, 107008, 107040]
    for a, b in zip(offs[:-1], offs[1:]):
        d = np.linalg.norm(g[a:b] - og[a:b]) / max(np.linalg.norm(og[a:b]), 1e-30)
        assert d < 3e-2, (a, d)
    cos = np.dot(g, og) / (np.linalg.norm(g) * np.linalg.norm(og))
    assert cos > 0.999


def test_mlp_features_from_pm_grids(ct2):
    x, y, z = ct2.ic_plummer_mix(20000, 10000, 4, 0.02, 0.2, 8)
    n = x.size
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    gm = float(np.float32(1.0 / n))
    ct2.set_force_options(1e-3, gm, 0)
    ng = 32
    ct2.pm_init(ng, 1.25)
    ct2.mlp_init(1)
    ct2.pm_forces()
    W = ct2.pm_density_download().astype(np.float64) * 2.0 ** -32
    A = ct2.pm_grid_download(0)
    B = ct2.pm_grid_download(1)
    idx = np.array([0, 5, 999, 15000, 19999])
    X = ct2.mlp_features(idx)
    s = 32 - 5
    funit = gm * ng * ng
    for r, p in enumerate(idx):
        c = [((int(v[p]) + (1 << (s - 1))) >> s) % ng for v in (x, y, z)]
        # the wrapped 32-bit difference, as the kernel takes it
        off = [np.float32(np.uint32((int(v[p]) - (cc << s)) & 0xFFFFFFFF).view(np.int32)) / np.float32(2 ** s)
               for v, cc in zip((x, y, z), c)]
        assert np.allclose(X[r, 108:111], off, atol=1e-7)
        q = 0
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for cz in (-1, 0, 1):
                    i, j, k = (c[0] + a) % ng, (c[1] + b) % ng, (c[2] + cz) % ng
                    assert X[r, 4 * q] == pytest.approx(W[i, j, k] * ng ** 3 / n, rel=1e-6, abs=1e-6)
                    assert X[r, 4 * q + 1] == pytest.approx(A[i, j, k].real / funit, rel=1e-5, abs=1e-6)
                    assert X[r, 4 * q + 2] == pytest.approx(A[i, j, k].imag / funit, rel=1e-5, abs=1e-6)
                    assert X[r, 4 * q + 3] == pytest.approx(B[i, j, k].real / funit, rel=1e-5, abs=1e-6)
                    q += 1
        assert not X[r, 111:].any()


def rel_err(F, J):
    return np.linalg.norm(F - J, axis=0) / np.linalg.norm(J, axis=0)


def test_mlp_training_reduces_pm_error(ct2):
    """Train on direct-sum (periodic FP64) labels of two lattice+Gaussian
    sets (0.2 cell, as the GPU configuration), then on held-out particles of
    a third seed the PM + MLP force must be closer to the judge than the PM
    force alone."""
    side, ng = 32, 64
    n = side ** 3
    gm = float(np.float32(1.0 / n))
    ct2.set_force_options(1.0 / (50 * side), gm, 0)
    ct2.pm_init(ng, 1.25)
    ct2.mlp_init(3)
    ct2.mlp_data_reset(1 << 16)
    for seed in (1, 2):
        x, y, z = ct2.ic_lattice(side, 0.2, seed)
        ct2.particle_allocate(n)
        ct2.particles_upload(x, y, z)
        ct2.mlp_data_append(np.random.default_rng(seed).choice(n, 16384, replace=False))
    assert ct2.mlp_data_size() == 32768
    losses = ct2.mlp_train(2000, 1024, 1e-3, 5)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-100:]) < 0.7 * np.mean(losses[:20])
    # held-out set
    x, y, z = ct2.ic_lattice(side, 0.2, 99)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    idx = np.random.default_rng(9).choice(n, 2048, replace=False)
    jx, jy, jz, _ = ct2.kick_sample_fp64(idx, periodic=True)
    J = np.stack([jx, jy, jz])
    fx, fy, fz, _ = ct2.pm_forces()
    e_pm = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J)
    fx, fy, fz, _ = ct2.pm_mlp_forces()
    e_ml = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J)
    print(f"held-out median rel err: PM {np.median(e_pm):.3e}  PM+MLP {np.median(e_ml):.3e}; "
          f"p90 {np.percentile(e_pm, 90):.3e} -> {np.percentile(e_ml, 90):.3e}", flush=True)
    assert np.median(e_ml) < 0.8 * np.median(e_pm)


def test_mlp_training_losses_are_deterministic_across_copy_chunks(ct2):
    """ct2_mlp_train keeps the step losses on the device and copies them in
    chunks of 4,096 steps: 4,100 steps fill every entry, and two runs from the
    same initial weights and seed give the same weights and losses bit for
    bit (fixed-order reductions throughout, no atomics on floats)."""
    side, ng = 16, 32
    n = side ** 3
    ct2.set_force_options(1.0 / (50 * side), float(np.float32(1.0 / n)), 0)
    ct2.pm_init(ng, 1.25)
    x, y, z = ct2.ic_lattice(side, 0.2, 4)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    ct2.mlp_init(7)
    ct2.mlp_data_reset(1024)
    ct2.mlp_data_append(np.random.default_rng(4).choice(n, 1024, replace=False))
    w_init = ct2.mlp_get_weights()
    runs = []
    for _ in range(2):
        ct2.mlp_set_weights(w_init)  # also clears Adam's moments and step count
        losses = ct2.mlp_train(4100, 128, 1e-3, 3)
        runs.append((losses, ct2.mlp_get_weights()))
    (l0, w0), (l1, w1) = runs
    assert np.isfinite(l0).all() and (l0 > 0).all()
    assert np.array_equal(l0, l1) and np.array_equal(w0, w1)
    assert np.mean(l0[-100:]) < np.mean(l0[:20])


def test_fused_inference_matches_gemm_forward(ct2, orc):
    """ct2_pm_mlp_forces runs the three layers in one kernel (transposed
    products, activations kept in registers, k-permuted weights).  Its
    correction must equal F_unit times the GEMM-path forward of the same
    features and the oracle's bf16 forward, to bf16 rounding; ragged n
    exercises the padded tail."""
    x, y, z = ct2.ic_plummer_mix(20011, 10000, 4, 0.02, 0.2, 12)
    n = x.size
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    gm = float(np.float32(1.0 / n))
    ct2.set_force_options(1e-3, gm, 0)
    ng = 32
    ct2.pm_init(ng, 1.25)
    ct2.mlp_init(5)
    w = ct2.mlp_get_weights()
    rng = np.random.default_rng(4)
    kin, hid, kout = 128, 256, 32
    b1 = kin * hid
    b2 = b1 + hid + hid * hid
    b3 = b2 + hid + hid * kout
    w[b1:b1 + hid] = rng.standard_normal(hid) * 0.3  # non-zero biases: exercise their indexing
    w[b2:b2 + hid] = rng.standard_normal(hid) * 0.3
    w[b3:b3 + 3] = rng.standard_normal(3) * 0.1
    ct2.mlp_set_weights(w)
    pm = np.stack(ct2.pm_forces()[:3]).astype(np.float64)
    fused = np.stack(ct2.pm_mlp_forces()[:3]).astype(np.float64)
    idx = np.concatenate([np.arange(0, n, 7), np.arange(n - 20, n)])
    X = ct2.mlp_features(idx)
    funit = gm * ng * ng
    got = ((fused - pm)[:, idx] / funit).T
    gemm = ct2.mlp_forward(X).astype(np.float64)
    want = orc.mlp_forward(w.astype(np.float64), X, bf16=True)[:, :3]
    scale = np.sqrt(np.mean(want ** 2))
    for ref in (gemm, want):
        err = np.abs(got - ref)
        assert np.median(err) < 2e-3 * scale and np.max(err) < 3e-2 * scale, (np.median(err) / scale, err.max() / scale)


def test_kick_pm_mlp_falls_back_to_the_direct_sum(ct2):
    """The direct-sum kick is the fallback of the learned correction:
    ct2_kick_pm_mlp runs ct2_kick (bit for bit the same velocities and rungs)
    when no trained weights are loaded (ct2_mlp_init alone) or the mesh
    recorded for the weights is not the current one (split radius,
    assignment, near field, particles per node); with the recorded mesh it
    takes the learned path.  ct2_last_force_path reports which ran."""
    side = 16
    n = side ** 3
    x, y, z = ct2.ic_lattice(side, 0.2, 3)
    rng = np.random.default_rng(3)
    v0 = [rng.standard_normal(n).astype(np.float32) * 0.01 for _ in range(3)]
    params = ct2.kick_params(ascending=True, descending=True, top=True, rung=0, t0=0.01, eta=0.15, max_dt=0.01)

    def reset():
        ct2.particles_upload(x, y, z, *v0, np.zeros(n, np.int8))

    ct2.particle_allocate(n)
    ct2.set_force_options(float(np.float32(1.0 / (50 * side))), float(np.float32(1.0 / n)), 0)
    reset()
    want_max = ct2.kick(params, ct2.KICK_STORE)
    assert ct2.last_force_path() == ct2.PATH_DIRECT
    want = ct2.particles_download()
    try:
        ct2.pm_assignment(2)
        ct2.pm_init(2 * side, 1.0)
        ct2.mlp_init(4)  # random weights, never trained
        for setup in ("untrained", "other split", "other density", "other assignment"):
            if setup == "other split":
                ct2.mlp_set_weights(ct2.mlp_get_weights())
                ct2.mlp_set_mesh(1.25, 2, 0.0, 1.0 / 8)
            elif setup == "other density":
                ct2.mlp_set_mesh(1.0, 2, 0.0, 1.0)
            elif setup == "other assignment":
                ct2.mlp_set_mesh(1.0, 1, 0.0, 1.0 / 8)
            reset()
            mr = ct2.kick_pm_mlp(params, ct2.KICK_STORE)
            got = ct2.particles_download()
            assert ct2.last_force_path() == ct2.PATH_DIRECT, setup
            assert mr == want_max
            for a in ("vx", "vy", "vz", "rung"):
                assert np.array_equal(getattr(got, a), getattr(want, a)), (setup, a)
        ct2.mlp_set_mesh(1.0, 2, 0.0, 1.0 / 8)  # the mesh in use: 16^3 particles on 32^3 nodes
        reset()
        ct2.kick_pm_mlp(params, ct2.KICK_STORE)
        assert ct2.last_force_path() == ct2.PATH_PM_MLP
        ct2.kick_pm(params, ct2.KICK_STORE)
        assert ct2.last_force_path() == ct2.PATH_PM
    finally:
        ct2.pm_assignment(1)
