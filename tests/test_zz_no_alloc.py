"""ct2_kick_noalloc, the kick that allocates nothing, against ct2_kick with
CT2_KICK_ORDERED through the C ABI: two calls in a row on the same
particles, the second on a smaller active range, each compared on forces,
potentials, the epilogue's velocities and rungs and the returned max rung;
bit for bit where both split the sources alike (they do at these sizes).
Every library allocation is made to fail around the no-allocation calls,
and an active range beyond the maximum is refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KICK_TILE = 256


def source_chunk(ns, nsrc, cap_bytes=None):
    """The source chunk the ordered pass picks (its split rule, and the
    reduction of the split to a fixed buffer)."""
    nb = -(-ns // 256)
    nsplit = min(max(1, nsrc // (4 * KICK_TILE)), max(1, -(-4096 // nb)))
    if cap_bytes is not None:
        nsplit = max(1, min(nsplit, cap_bytes // (32 * ns)))
    jchunk = -(-nsrc // nsplit)
    return -(-jchunk // KICK_TILE) * KICK_TILE


def state(ct2):
    p = ct2.particles_download()
    return p, ct2.forces_download()


def compare(ct2, ref, got, ns_first, bit_exact):
    (p0, f0), (p1, f1) = ref, got
    for a, b in zip(f0, f1):
        if bit_exact:
            assert np.array_equal(a[ns_first:], b[ns_first:])
        else:
            assert np.allclose(a[ns_first:], b[ns_first:], rtol=2e-6, atol=0)
    for name in ("vx", "vy", "vz", "rung", "x", "y", "z"):
        a, b = getattr(p0, name), getattr(p1, name)
        assert np.array_equal(a, b) if bit_exact or name in ("x", "y", "z") else np.allclose(a, b, rtol=2e-6), name


def test_kick_noalloc_matches_ordered_kick(ct2):
    n = 20000
    x, y, z = ct2.ic_plummer_mix(n, 8000, 4, 0.02, 0.2, 23)
    rng = np.random.default_rng(5)
    rung = rng.integers(0, 4, n).astype(np.int8)
    v = [(rng.standard_normal(n) * 0.05).astype(np.float32) for _ in range(3)]
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, *v, rung)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(2e-3))
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    params = ct2.kick_params(ascending=True, descending=True, top=True, first=False, rung=1, t0=0.02, eta=0.15,
                             max_dt=1.0)
    cap = 32 * ct2.NOALLOC_MAX_SINKS
    ct2.particle_sort_by_rung(1)
    ns_prev = n
    for call, minrung in enumerate((1, 3)):  # the second call: a smaller active range
        if call:  # fresh rungs (the first kick's epilogue moved them), then a narrower range
            ct2.particles_upload(None, None, None, None, None, None, rng.integers(0, 4, n).astype(np.int8))
            ct2.particle_push_rungs()
            ct2.particle_sort_by_rung(minrung)
        first, last = ct2.particle_current_range()
        ns = n - first
        assert 1000 < ns < ns_prev
        ns_prev = ns
        start = ct2.particles_download()
        r_ref = ct2.kick(params, ct2.KICK_ORDERED | ct2.KICK_STORE)
        ref = state(ct2)
        ct2.particles_upload(start.x, start.y, start.z, start.vx, start.vy, start.vz, start.rung)
        ct2.debug_fail_allocations(0, 1000)  # any allocation now fails with CT2_ERR_NOMEM
        try:
            r_got = ct2.kick_noalloc(params, ct2.KICK_STORE)
        finally:
            ct2.debug_fail_allocations(0, 0)
        got = state(ct2)
        assert ct2.last_force_path() == ct2.PATH_DIRECT
        bit_exact = source_chunk(ns, n) == source_chunk(ns, n, cap)
        assert bit_exact  # these sizes split the sources alike
        assert r_got == r_ref
        compare(ct2, ref, got, first, bit_exact)
        assert not np.array_equal(got[0].vx[first:], start.vx[first:])  # the epilogue ran
        assert np.array_equal(got[0].vx[:first], start.vx[:first])  # and only on the range
    ct2.particle_pop_rungs()


def test_kick_noalloc_refuses_beyond_its_maximum(ct2):
    from ct2_amd._lib import ERR_INVALID, Ct2Error

    n = ct2.NOALLOC_MAX_SINKS + 1
    ct2.particle_allocate(n)
    rng = np.random.default_rng(1)
    x, y, z = (rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    ct2.particles_upload(x, y, z, np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32),
                         np.zeros(n, np.int8))
    ct2.set_force_options(0.01, 1.0 / n, ct2.KERNEL_PLUMMER)
    before = ct2.particles_download()
    with pytest.raises(Ct2Error) as e:
        ct2.kick_noalloc(ct2.kick_params(ascending=True, top=True, rung=0, t0=0.01))
    assert e.value.code == ERR_INVALID and "maximum" in str(e.value)
    with pytest.raises(Ct2Error) as e:
        ct2.kick_noalloc(None, ct2.KICK_SYMMETRIC)
    assert e.value.code == ERR_INVALID
    after = ct2.particles_download()
    assert np.array_equal(before.vx, after.vx)
    # exactly the maximum: the largest range it takes, summed in one source split
    ct2.particle_allocate(ct2.NOALLOC_MAX_SINKS)
    ct2.particles_upload(x[:-1], y[:-1], z[:-1], np.zeros(n - 1, np.float32), np.zeros(n - 1, np.float32),
                         np.zeros(n - 1, np.float32), np.zeros(n - 1, np.int8))
    ct2.kick_noalloc(None, ct2.KICK_STORE)
    f = ct2.forces_download()
    ct2.kick(None, ct2.KICK_ORDERED | ct2.KICK_STORE)
    g = ct2.forces_download()
    for a, b in zip(f, g):
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(b).max())
