"""The multipole-mesh force (particle.cpp:116-271, kick.cpp:11-29, 105-125) on
the GPU against the FP64 oracle's restatement: the derivative table, the
force and potential (near-set pairs + far field) on uniform and clustered
sets at several orders, the kick epilogue on the current rung range, and at
the lattice size of the GPU configuration against the periodic FP64 judge."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2 ** 32


def to_fixed(p):
    return np.round((np.asarray(p) % 1.0) * TWO32).astype(np.uint64).astype(np.uint32)


def random_set(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 3))
    if clustered:
        c = rng.random((4, 3))
        k = n // 2
        p[:k] = c[rng.integers(0, 4, k)] + 0.03 * rng.standard_normal((k, 3))
    return [to_fixed(p[:, q]) for q in range(3)]


def rel_err(F, W):
    return np.linalg.norm(F - W, axis=0) / np.linalg.norm(W, axis=0)


@pytest.mark.parametrize("nc,order", [(8, 4), (16, 8), (32, 2)])
def test_table_matches_oracle(ct2, orc, nc, order):
    ct2.particle_allocate(16)
    ct2.mm_init(nc, order)
    got = ct2.mm_table_download()
    want = orc.mm_green_table(order, nc)
    scale = np.abs(want).reshape(want.shape[0], -1).max(1)
    err = np.abs(got - want).reshape(want.shape[0], -1).max(1) / scale
    assert err.max() < 1e-11, err.max()


@pytest.mark.parametrize("nc,order", [(8, 1), (8, 4), (8, 8), (16, 4), (16, 8), (32, 6)])
@pytest.mark.parametrize("clustered", [False, True])
def test_forces_vs_oracle(ct2, orc, nc, order, clustered):
    """8^3 cells take the exact-separation path, 16^3 and 32^3 the FP32-offset
    path with per-sink runs (clustered: columns of several hundred sources,
    staged in more than one pass, and windows wrapping through z = 0)."""
    n = 3000
    x, y, z = random_set(n, 21 + order, clustered)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(2e-3))
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.mm_init(nc, order)
    fx, fy, fz, pot = ct2.mm_forces()
    assert ct2.last_force_path() == ct2.PATH_MM
    ox, oy, oz, op = orc.mm_forces(x, y, z, nc, order, gm, soft)
    e = rel_err(np.stack([fx, fy, fz]), np.stack([ox, oy, oz]))
    print(f"{nc}^3 cells order {order} clustered {clustered}: GPU vs oracle median {np.median(e):.2e} max {e.max():.2e}",
          flush=True)
    assert np.median(e) < 3e-6 and np.percentile(e, 99) < 5e-5, (np.median(e), e.max())
    assert np.max(np.abs(pot - op)) < 1e-5 * np.abs(op).max()  # the periodic potential crosses zero


def test_ct2_kernel_and_empty_cells(ct2, orc):
    """ct2's uniform-sphere kernel (kick.cpp:52-60) in the direct pairs, a
    16^3 grid with most cells empty, and particles on cell faces."""
    nc, n = 16, 700
    x, y, z = random_set(n, 5, clustered=True)
    x[:10] = (np.arange(10, dtype=np.uint64) << 28).astype(np.uint32)  # on x cell faces
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(1e-2))
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(soft, gm, ct2.KERNEL_CT2)
    try:
        ct2.mm_init(nc, 5)
        fx, fy, fz, pot = ct2.mm_forces()
        ox, oy, oz, op = orc.mm_forces(x, y, z, nc, 5, gm, soft, kernel=orc.KERNEL_CT2)
    finally:
        ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    e = rel_err(np.stack([fx, fy, fz]), np.stack([ox, oy, oz]))
    assert np.median(e) < 3e-6 and np.percentile(e, 99) < 5e-5, (np.median(e), e.max())
    assert np.max(np.abs(pot - op)) < 1e-5 * np.abs(op).max()  # the periodic potential crosses zero


def test_dense_cluster_across_the_z_boundary(ct2, orc):
    """A cluster of 2,500 particles (sigma 0.008 box) centred on a z cell face
    at z = 0, on 16^3 cells: its columns hold windows of over a thousand
    sources (staged in several passes) that wrap through z = 0."""
    nc, n = 16, 4000
    rng = np.random.default_rng(31)
    p = rng.random((n, 3))
    p[:2500] = np.array([0.40, 0.55, 0.0]) + 0.008 * rng.standard_normal((2500, 3))
    x, y, z = [to_fixed(p[:, q]) for q in range(3)]
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(1e-3))
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.mm_init(nc, 4)
    fx, fy, fz, pot = ct2.mm_forces()
    ox, oy, oz, op = orc.mm_forces(x, y, z, nc, 4, gm, soft)
    e = rel_err(np.stack([fx, fy, fz]), np.stack([ox, oy, oz]))
    print(f"dense cluster: GPU vs oracle median {np.median(e):.2e} max {e.max():.2e}", flush=True)
    assert np.median(e) < 3e-6 and np.percentile(e, 99) < 5e-5, (np.median(e), e.max())
    assert np.max(np.abs(pot - op)) < 1e-5 * np.abs(op).max()


@pytest.mark.parametrize("nc,n", [(8, 1), (8, 5), (16, 2), (16, 37), (32, 64)])
def test_few_particles(ct2, orc, nc, n):
    """Ragged and tiny sets: one particle (its force is the periodic self term,
    zero by symmetry, up to the expansion's truncation), a handful, and grids
    with far more cells than particles."""
    x, y, z = random_set(n, 100 + n)
    gm, soft = float(np.float32(1.0 / max(n, 2))), float(np.float32(2e-3))
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.mm_init(nc, 4)
    fx, fy, fz, pot = ct2.mm_forces()
    ox, oy, oz, op = orc.mm_forces(x, y, z, nc, 4, gm, soft)
    F, W = np.stack([fx, fy, fz]).astype(np.float64), np.stack([ox, oy, oz])
    # one particle: its force is zero, measured against a pair one cell apart (GM nc^2)
    scale = np.abs(W).max() if n > 1 else gm * nc * nc
    assert np.abs(F - W).max() < 1e-5 * scale, (np.abs(F - W).max(), scale)
    assert np.max(np.abs(pot - op)) < 1e-5 * max(np.abs(op).max(), gm)


def test_kick_mm_applies_the_epilogue_to_the_range(ct2, orc):
    n = 4000
    x, y, z = random_set(n, 8)
    rng = np.random.default_rng(2)
    rung = rng.integers(0, 3, n).astype(np.int8)
    v = [(0.01 * rng.standard_normal(n)).astype(np.float32) for _ in range(3)]
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, *v, rung)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(2e-3))
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.particle_sort_by_rung(1)
    first, _ = ct2.particle_current_range()
    p0 = ct2.particles_download()
    ct2.mm_init(8, 4)
    params = ct2.kick_params(ascending=True, descending=True, top=True, rung=1, t0=0.01)
    mr = ct2.kick_mm(params, ct2.KICK_STORE)
    f = ct2.forces_download()
    p1 = ct2.particles_download()
    assert np.array_equal(p1.vx[:first], p0.vx[:first]) and np.array_equal(p1.rung[:first], p0.rung[:first])
    vx, vy, vz, rg = (p0.vx[first:].copy(), p0.vy[first:].copy(), p0.vz[first:].copy(), p0.rung[first:].copy())
    op = orc.KickParams(*[getattr(params, k) for k, _ in params._fields_])
    omr = orc.kick_epilogue(f[0][first:], f[1][first:], f[2][first:], vx, vy, vz, rg, op, soft)
    assert np.array_equal(p1.vx[first:], vx) and np.array_equal(p1.rung[first:], rg)
    assert mr == omr


def test_converges_to_the_periodic_judge(ct2):
    """Lattice + Gaussian(0.2 cell) at 32^3 particles on 8^3 cells (64 per
    cell): against the FP64 Ewald direct sum the median force error falls
    with the order, to below 2e-5 at order 8."""
    ct2.particle_allocate(32 ** 3)
    x, y, z = ct2.ic_lattice(32, 0.2, 3)
    n = x.size
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(float(np.float32(1.0 / 1600)), float(np.float32(1.0 / n)), ct2.KERNEL_PLUMMER)
    idx = np.random.default_rng(1).choice(n, 512, replace=False)
    jx, jy, jz, _ = ct2.kick_sample_fp64(idx, periodic=True)
    J = np.stack([jx, jy, jz])
    med = []
    for order in (2, 4, 6, 8):
        ct2.mm_init(8, order)
        fx, fy, fz, _ = ct2.mm_forces()
        med.append(float(np.median(rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), J))))
    print("lattice 32^3, 8^3 cells: median error by order 2,4,6,8:", med, flush=True)
    assert all(b < 0.5 * a for a, b in zip(med, med[1:])), med
    assert med[-1] < 2e-5


def test_config2_lattice_against_the_judge(ct2):
    """N = 128^3 lattice + Gaussian(0.2 cell), 64^3 cells of 8 particles,
    order 8: median force error against the periodic FP64 judge on 4096
    sinks below 4e-5 (3.1e-5 measured; 1.7e-5 on 32^3 cells, where the
    direct pairs carry more of the force).  The potentials differ by a near-constant: the judge
    softens every pair's nearest image, the mesh only the direct pairs, so
    the far pairs add GM sum eps^2 / (2 r^3), about GM N 2 pi eps^2
    ln(r_max / r_near) ~ 0.5 GM here; about that offset the spread is below
    1e-5 of the potential."""
    ct2.particle_allocate(128 ** 3)
    x, y, z = ct2.ic_lattice(128, 0.2, 1)
    n = x.size
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(float(np.float32(1.0 / 6400)), float(np.float32(1.0 / n)), ct2.KERNEL_PLUMMER)
    ct2.mm_init(ct2.mm_cells_for(n), 8)
    fx, fy, fz, pot = ct2.mm_forces()
    idx = np.random.default_rng(42).choice(n, 4096, replace=False)
    jx, jy, jz, jp = ct2.kick_sample_fp64(idx, periodic=True)
    e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), np.stack([jx, jy, jz]))
    print(f"cfg2 multipole mesh (64^3 cells, order 8) vs periodic FP64: median {np.median(e):.2e} "
          f"p90 {np.percentile(e, 90):.2e} p99 {np.percentile(e, 99):.2e}", flush=True)
    assert np.median(e) < 4e-5
    d = pot[idx] - jp
    gm = float(np.float32(1.0 / n))
    print(f"potential offset {d.mean() / gm:.3f} GM, spread {d.std() / np.abs(jp).mean():.2e}", flush=True)
    assert 0.2 * gm < d.mean() < 1.0 * gm
    assert d.std() < 1e-5 * np.abs(jp).mean()
