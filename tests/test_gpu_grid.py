"""ct2's cell binning of the container (particle_to_grid, particle.cpp:59-114,
349-362) through the C ABI, and the failure path of a device allocation
(ct2 aborts, memory.cpp:7-10; the library reports CT2_ERR_NOMEM and stays
usable)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2 ** 32


def cell_of(x, y, z, nc):
    x, y, z = (v.astype(np.uint64) for v in (x, y, z))
    return ((x * nc >> 32) * nc + (y * nc >> 32)) * nc + (z * nc >> 32)


@pytest.mark.parametrize("n,nc,minrung", [(1, 1, 0), (1000, 3, 0), (65543, 16, 2), (300000, 37, 1)])
def test_to_grid_is_a_stable_cell_sort(ct2, n, nc, minrung):
    rng = np.random.default_rng(n + nc)
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    edge = np.array([0, TWO32 - 1, TWO32 // 2, TWO32 // 2 - 1], np.uint32)
    x[: min(n, 4)] = edge[: min(n, 4)]
    vx = np.arange(n, dtype=np.float32)  # tags (exact below 2^24)
    vy = rng.standard_normal(n).astype(np.float32)
    vz = rng.standard_normal(n).astype(np.float32)
    rung = rng.integers(0, 4, n).astype(np.int8)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, vx, vy, vz, rung)
    ct2.particle_sort_by_rung(minrung)
    first, last = ct2.particle_current_range()
    p0 = ct2.particles_download()
    starts = ct2.particles_to_grid(nc)
    p = ct2.particles_download()
    # outside the range nothing moves
    for a in ("x", "y", "z", "vx", "vy", "vz", "rung"):
        assert np.array_equal(getattr(p, a)[:first], getattr(p0, a)[:first])
    key0 = cell_of(p0.x[first:], p0.y[first:], p0.z[first:], nc)
    order = np.argsort(key0, kind="stable")
    for a in ("x", "y", "z", "vx", "vy", "vz", "rung"):
        assert np.array_equal(getattr(p, a)[first:], getattr(p0, a)[first:][order]), a
    key = key0[order]
    cells = np.arange(nc ** 3 + 1)
    assert np.array_equal(starts, first + np.searchsorted(key, cells, side="left"))
    assert starts[-1] == last


def test_to_grid_rejects_bad_sizes(ct2):
    from ct2_amd import Ct2Error
    ct2.particle_allocate(10)
    for nc in (0, -1, 1025):
        with pytest.raises(Ct2Error):
            ct2.particles_to_grid(nc)


def test_injected_allocation_failure(ct2):
    """ct2_debug_fail_allocations(after, count) makes device allocations fail
    without calling hipMalloc: the call reports CT2_ERR_NOMEM naming the
    allocation, leaves no half-made container (a failure at any of the
    container's eleven arrays), and the library works once the hook is spent,
    also when the failing allocation is a kernel's scratch buffer."""
    from ct2_amd import Ct2Error
    from ct2_amd import _lib as L
    for after in (0, 4, 10):
        ct2.debug_fail_allocations(after, 1)
        with pytest.raises(Ct2Error) as e:
            ct2.particle_allocate(4096)
        assert e.value.code == L.ERR_NOMEM and "test hook" in str(e.value)
        assert ct2.particle_count() == 0
    ct2.debug_fail_allocations(0, 0)
    # a fresh library state, so that the kick has to allocate its partial sums
    ct2.finalize()
    ct2.init(0)
    x, y, z = ct2.ic_lattice(16, 0.2, 1)
    ct2.particle_allocate(x.size)
    ct2.particles_upload(x, y, z)
    ct2.set_force_options(1e-3, 1.0 / x.size, 0)
    ct2.debug_fail_allocations(0, 1)
    with pytest.raises(Ct2Error) as e:
        ct2.kick(None, ct2.KICK_STORE)
    assert e.value.code == L.ERR_NOMEM
    ct2.kick(None, ct2.KICK_STORE)  # the hook is spent
    fx, fy, fz, pot = ct2.forces_download()
    assert np.isfinite(fx).all() and (pot > 0).all()
