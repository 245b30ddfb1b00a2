"""GPU parity of the kick and the drift against the FP64 CPU oracle, through
the C ABI (kick.cpp:36-99, drift.cpp:4-17)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2 ** 32


def f32(v):
    """The library keeps hsoft and GM as FP32 options (options.hpp:20-26);
    the oracle gets the same rounded values."""
    return float(np.float32(v))


def rel_err(a, b):
    """per-particle |a-b|/|b| for (3, n) force arrays"""
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


def load(ct2, x, y, z, vx=None, vy=None, vz=None, rung=None):
    ct2.particle_allocate(x.size)
    ct2.particles_upload(x, y, z, vx, vy, vz, rung)


# ------------------------------------------------------------------ drift
@pytest.mark.parametrize("n", [1, 3, 4, 1001, 65536 + 7])
def test_drift_bit_exact(ct2, orc, n):
    rng = np.random.default_rng(n)
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    vx = (rng.standard_normal(n) * 0.3).astype(np.float32)
    vy = (rng.standard_normal(n) * 30).astype(np.float32)
    vz = (rng.standard_normal(n) * 1e-6).astype(np.float32)
    edge = np.array([0.25, -0.25, 2.0 ** -33, 3 * 2.0 ** -33, 0.75, -0.75, 1.25, 3e9, np.nan, np.inf, -0.0],
                    np.float32)
    vx[: min(n, edge.size)] = edge[: min(n, edge.size)]
    load(ct2, x, y, z, vx, vy, vz)
    ct2.drift(0.37, 1.3)
    got = ct2.particles_download()
    want = orc.drift(x, y, z, vx, vy, vz, 0.37, 1.3)
    for g, w in zip((got.x, got.y, got.z), want):
        assert np.array_equal(g, w)
    ct2.drift(-0.37, 1.3)  # time reversal returns the start exactly
    back = ct2.particles_download()
    assert np.array_equal(back.x, x) and np.array_equal(back.y, y) and np.array_equal(back.z, z)


# ------------------------------------------------------------------ kick
def lattice(ct2, n, sigma, seed):
    return ct2.ic_lattice(n, sigma, seed)


def test_pinned_host_arrays_round_trip(ct2):
    """The container and the stored forces through page-locked host arrays
    (ct2_host_alloc): an upload and download give the data back bit for bit,
    and the force download into them equals the one into new arrays."""
    n = 4099
    rng = np.random.default_rng(17)
    src = [rng.integers(0, 2 ** 32, n, dtype=np.uint32) for _ in range(3)]
    src += [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    src.append(rng.integers(0, 8, n).astype(np.int8))
    pin = [ct2.pinned_empty(n, a.dtype) for a in src]
    for d, a in zip(pin, src):
        d[:] = a
    ct2.particle_allocate(n)
    ct2.particles_upload(*pin)
    back = ct2.particles_download()
    for a, b in zip(src, (back.x, back.y, back.z, back.vx, back.vy, back.vz, back.rung)):
        assert np.array_equal(a, b)
    ct2.set_force_options(1e-3, float(np.float32(1.0 / n)), ct2.KERNEL_PLUMMER)
    ct2.kick(ct2.kick_params(rung=0, t0=0.0), ct2.KICK_STORE)
    want = ct2.forces_download()
    got = ct2.forces_download([ct2.pinned_empty(n, np.float32) for _ in range(4)])
    for a, b in zip(want, got):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("side", [16, 20])
def test_kick_all_sinks_vs_fp64(ct2, orc, kernel, side):
    """Small sets take the split-source path (several source chunks per sink
    block, reduced in FP64 in a fixed order)."""
    x, y, z = lattice(ct2, side, 0.2, 11)
    n = x.size
    soft, gm = f32(1.0 / (50 * side)), f32(1.0 / n)
    load(ct2, x, y, z)
    ct2.set_force_options(soft, gm, kernel)
    ct2.kick(None, ct2.KICK_STORE)
    fx, fy, fz, pot = ct2.forces_download()
    ax, ay, az, ap = orc.kick(x, y, z, gm, soft, kernel)
    e = rel_err(np.stack([fx, fy, fz]), np.stack([ax, ay, az]))
    assert np.median(e) < 2e-6 and np.percentile(e, 99) < 5e-5, (np.median(e), e.max())
    assert np.max(np.abs(pot / ap - 1)) < 1e-6


def test_kick_fp64_judge_matches_cpu_oracle(ct2, orc):
    x, y, z = ct2.ic_plummer_mix(20000, 10000, 8, 0.01, 0.2, 5)
    load(ct2, x, y, z)
    for kernel in (0, 1):
        ct2.set_force_options(2e-4, 1.0 / x.size, kernel)
        idx = np.arange(0, x.size, 97)
        g = ct2.kick_sample_fp64(idx)
        c = orc.kick(x, y, z, f32(1.0 / x.size), f32(2e-4), kernel, sinks=idx)
        # FP64 sums of the same terms in another order
        for a, b in zip(g, c):
            assert np.allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(b).max())


def test_kick_epilogue_bit_exact(ct2, orc):
    """The fused half-kick/rung epilogue equals the CPU restatement of
    kick.cpp:75-99 bit for bit when given the same FP32 forces."""
    x, y, z = ct2.ic_plummer_mix(5000, 2500, 4, 0.02, 0.3, 9)
    n = x.size
    rng = np.random.default_rng(1)
    vx, vy, vz = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    rung = rng.integers(0, 6, n).astype(np.int8)
    load(ct2, x, y, z, vx, vy, vz, rung)
    ct2.set_force_options(1e-3, 1.0 / n, 0)
    for p in (ct2.kick_params(ascending=1, descending=1, top=1, first=0, rung=2, t0=0.5, eta=0.2, max_dt=0.1, scale=0.7),
              ct2.kick_params(ascending=1, descending=0, top=0, first=0, rung=3, t0=0.25),
              ct2.kick_params(ascending=0, descending=1, top=1, first=1, rung=0, t0=1.0, eta=0.05, max_dt=0.5)):
        before = ct2.particles_download()
        mr = ct2.kick(p, ct2.KICK_STORE)
        fx, fy, fz, _ = ct2.forces_download()
        after = ct2.particles_download()
        ov, ow, ou, orr = before.vx.copy(), before.vy.copy(), before.vz.copy(), before.rung.copy()
        omr = orc.kick_epilogue(fx, fy, fz, ov, ow, ou, orr, orc.KickParams(*[getattr(p, f) for f, _ in p._fields_]),
                                1e-3)
        assert np.array_equal(after.vx, ov) and np.array_equal(after.vy, ow) and np.array_equal(after.vz, ou)
        assert np.array_equal(after.rung, orr)
        assert mr == omr


def test_kick_config2_full_size_vs_fp64_judge(ct2, orc):
    """N = 128^3 lattice + Gaussian(0.2 cell), eps = 1/(50*128): FP32
    compensated kick against the FP64 direct sum on 4096 seeded sinks (GPU
    judge) and 128 of them on the CPU oracle."""
    x, y, z = lattice(ct2, 128, 0.2, 1)
    n = x.size
    soft, gm = f32(1.0 / (50 * 128)), f32(1.0 / n)
    load(ct2, x, y, z)
    ct2.set_force_options(soft, gm, 0)
    ct2.kick(None, ct2.KICK_STORE)
    fx, fy, fz, pot = ct2.forces_download()
    idx = np.random.default_rng(42).choice(n, 4096, replace=False)
    jx, jy, jz, jp = ct2.kick_sample_fp64(idx)
    e = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), np.stack([jx, jy, jz]))
    print(f"cfg2 FP32 kick vs FP64: median {np.median(e):.3e} p90 {np.percentile(e, 90):.3e} "
          f"p99 {np.percentile(e, 99):.3e}", flush=True)
    assert np.median(e) < 1e-5 and np.percentile(e, 99) < 1e-3
    assert np.max(np.abs(pot[idx] / jp - 1)) < 1e-5
    sub = idx[:128]
    cx, cy, cz, cp = orc.kick(x, y, z, gm, soft, 0, sinks=sub)
    assert np.allclose(jx[:128], cx, rtol=1e-9, atol=1e-11 * np.abs(cx).max())
    assert np.allclose(jp[:128], cp, rtol=1e-12)


def test_kahan_beats_naive_on_clustered_set(ct2):
    """Precision study (clustered inputs): compensated accumulation is at
    least as accurate as naive FP32 at every percentile."""
    n = 1 << 18
    x, y, z = ct2.ic_plummer_mix(n, n // 2, 64, 0.01, 0.1, 3)
    load(ct2, x, y, z)
    ct2.set_force_options(2e-4, 1.0 / n, 0)
    idx = np.random.default_rng(3).choice(n, 2048, replace=False)
    jx, jy, jz, jp = ct2.kick_sample_fp64(idx)
    ref = np.stack([jx, jy, jz])
    errs = {}
    for name, flag in (("kahan", 0), ("naive", ct2.KICK_NAIVE)):
        ct2.kick(None, flag | ct2.KICK_STORE)
        fx, fy, fz, pot = ct2.forces_download()
        errs[name] = rel_err(np.stack([fx[idx], fy[idx], fz[idx]]), ref)
        errs[name + "_pot"] = np.abs(pot[idx] / jp - 1)
    for q in (50, 90, 99):
        assert np.percentile(errs["kahan"], q) <= np.percentile(errs["naive"], q) * 1.05
    assert np.median(errs["kahan_pot"]) < np.median(errs["naive_pot"])


@pytest.mark.parametrize("n", [1, 2, 255, 257, 1000])
def test_kick_ragged_sizes(ct2, orc, n):
    rng = np.random.default_rng(n)
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    load(ct2, x, y, z)
    ct2.set_force_options(0.01, 1.0 / n, 0)
    ct2.kick(None, ct2.KICK_STORE)
    fx, fy, fz, pot = ct2.forces_download()
    ax, ay, az, ap = orc.kick(x, y, z, f32(1.0 / n), f32(0.01), 0)
    if n > 1:
        e = rel_err(np.stack([fx, fy, fz]), np.stack([ax, ay, az]))
        assert np.max(e) < 1e-4
    else:
        assert fx[0] == 0 and fy[0] == 0 and fz[0] == 0
    assert np.allclose(pot, ap, rtol=1e-6)


@pytest.mark.parametrize("n", [1024, 1025, 2048, 3001, 3072, 4096, 65536, 65543])
@pytest.mark.parametrize("kernel", [0, 1])
def test_symmetric_and_ordered_kicks_vs_fp64(ct2, orc, n, kernel):
    """Every particle a sink: the symmetric kernels evaluate each pair of the
    last floor(n / 1024) * 1024 particles once and apply it to both (one
    superblock; two; odd and even superblock counts; 64 superblocks = two
    launches of 16 rounds); for ragged n the ordered kernel adds the pairs of
    that core with the first n % 1024 particles and those particles' own
    sums.  CT2_KICK_SYMMETRIC forces that split at these sizes (by default a
    core below 2^16 particles is left to the ordered kernel); the forced,
    ordered and default paths against the FP64 oracle, on a clustered set."""
    x, y, z = ct2.ic_plummer_mix(n, n // 2, 4, 0.02, 0.2, n + kernel)
    load(ct2, x, y, z)
    soft, gm = f32(2e-3), f32(1.0 / n)
    ct2.set_force_options(soft, gm, kernel)
    sinks = np.arange(n) if n <= 4096 else np.random.default_rng(1).choice(n, 1024, replace=False)
    ax, ay, az, ap = orc.kick(x, y, z, gm, soft, kernel, sinks=sinks)
    A = np.stack([ax, ay, az])
    for flag in (ct2.KICK_SYMMETRIC, ct2.KICK_ORDERED, 0):
        ct2.kick(None, flag | ct2.KICK_STORE)
        fx, fy, fz, pot = ct2.forces_download()
        e = rel_err(np.stack([fx[sinks], fy[sinks], fz[sinks]]), A)
        assert np.median(e) < 2e-6 and np.percentile(e, 99) < 5e-5, (flag, np.median(e), e.max())
        assert np.max(np.abs(pot[sinks] / ap - 1)) < 1e-6
        if flag == ct2.KICK_SYMMETRIC:  # Newton's third law: the pair forces cancel in the total
            F = np.stack([fx, fy, fz]).astype(np.float64)
            assert np.abs(F.sum(1)).max() < 1e-6 * np.abs(F).sum(1).max()


def test_kick_rejects_bad_params(ct2):
    from ct2_amd import Ct2Error
    ct2.particle_allocate(10)
    with pytest.raises(Ct2Error):
        ct2.kick(ct2.kick_params(descending=1, rung=31, t0=1.0))
    with pytest.raises(Ct2Error):
        ct2.kick(ct2.kick_params(descending=1, rung=0, t0=0.0))
    with pytest.raises(Ct2Error):
        ct2.kick_sample_fp64([10])
    with pytest.raises(Ct2Error):
        ct2.kick(None, ct2.KICK_ORDERED | ct2.KICK_SYMMETRIC)
