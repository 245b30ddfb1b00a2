"""Rung bookkeeping on the device (particle.cpp:20-58, 320-347) and the
kick_step schedule built on it (main.cpp:8-92): the partition, the range
stack, counts and minrung against numpy; kicks restricted to the current
range against the FP64 oracle; and the schedule the binding's kick_step runs,
replayed by an independent restatement of the reference's control flow."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_bin", "compat_driver")
TWO32 = 2 ** 32


def setup_rungs(ct2, n, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(n, dtype=np.uint32) * np.uint32(4099)  # distinct ids in x
    y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(2))
    rung = rng.integers(0, 6, n).astype(np.int8)
    vx = rng.standard_normal(n).astype(np.float32)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, vx, vx, vx, rung)
    return x, y, z, vx, rung


def test_sort_by_rung_is_a_stable_partition(ct2):
    n = 5000
    x, y, z, vx, rung = setup_rungs(ct2, n, 1)
    assert ct2.particle_current_range() == (0, n)
    ct2.particle_sort_by_rung(2)
    p = ct2.particles_download()
    order = np.concatenate([np.flatnonzero(rung < 2), np.flatnonzero(rung >= 2)])
    assert np.array_equal(p.x, x[order]) and np.array_equal(p.rung, rung[order])
    assert np.array_equal(p.y, y[order]) and np.array_equal(p.vx, vx[order])
    first = int((rung < 2).sum())
    assert ct2.particle_current_range() == (first, n)
    ct2.particle_push_rungs()
    ct2.particle_sort_by_rung(4)
    q = ct2.particles_download()
    sub = p.rung[first:]
    order2 = np.concatenate([np.arange(first), first + np.flatnonzero(sub < 4), first + np.flatnonzero(sub >= 4)])
    assert np.array_equal(q.x, p.x[order2])
    assert ct2.particle_current_range() == (first + int((sub < 4).sum()), n)
    ct2.particle_pop_rungs()
    assert ct2.particle_current_range() == (first, n)
    ct2.particle_sort_by_rung(0)
    assert ct2.particle_current_range() == (0, n)
    from ct2_amd._lib import Ct2Error
    with pytest.raises(Ct2Error):
        ct2.particle_pop_rungs()


def test_rung_counts_and_set_minrung(ct2):
    n = 7777
    _, _, _, _, rung = setup_rungs(ct2, n, 2)
    assert ct2.particle_rung_counts() == np.bincount(rung).tolist()
    ct2.particle_set_minrung(3)
    p = ct2.particles_download()
    assert np.array_equal(p.rung, np.maximum(rung, 3))
    assert ct2.particle_rung_counts() == np.bincount(np.maximum(rung, 3)).tolist()


def test_kicks_act_on_the_current_range(ct2, orc):
    """After sort_by_rung the kick's sinks are [first, n) and its sources all
    particles; particles below the range keep their velocities."""
    n = 3000
    rng = np.random.default_rng(3)
    x, y, z = (rng.integers(0, TWO32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    rung = rng.integers(0, 3, n).astype(np.int8)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, None, None, None, rung)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(0.01))
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.particle_sort_by_rung(1)
    first, _ = ct2.particle_current_range()
    p0 = ct2.particles_download()
    params = ct2.kick_params(ascending=True, descending=False, top=True, first=False, rung=1, t0=0.01)
    ct2.kick(params)
    p1 = ct2.particles_download()
    assert np.array_equal(p1.vx[:first], p0.vx[:first]) and np.array_equal(p1.rung, p0.rung)
    ax, ay, az, _ = orc.kick(p0.x, p0.y, p0.z, gm, soft, orc.KERNEL_PLUMMER, sinks=np.arange(first, n))
    dt = float(np.float32(0.5) * np.float32(0.01) / np.float32(2.0))  # 0.5 t0 / 2^rung
    W = dt * np.stack([ax, ay, az])
    V = np.stack([p1.vx[first:], p1.vy[first:], p1.vz[first:]]).astype(np.float64)
    err = np.linalg.norm(V - W, axis=0) / np.linalg.norm(W, axis=0)
    assert np.median(err) < 2e-6 and err.max() < 1e-4
    # the PM kick applies its epilogue to the same range
    ct2.pm_init(32, 1.25)
    ct2.kick_pm(params)
    p2 = ct2.particles_download()
    assert np.array_equal(p2.vx[:first], p0.vx[:first])
    assert not np.array_equal(p2.vx[first:], p1.vx[first:])


@pytest.mark.parametrize("flag_name", ["KICK_SYMMETRIC", "KICK_ORDERED"])
def test_subset_kick_with_symmetric_core(ct2, orc, flag_name):
    """Active range [first, n) of 3,000 + 2,500 particles: with
    CT2_KICK_SYMMETRIC its last 2,048 sinks form the symmetric core, the
    ordered kernel adds the core's pairs with everything before it and the
    other active sinks' sums; all against the FP64 oracle, and the inactive
    particles untouched."""
    n = 5500
    rng = np.random.default_rng(9)
    x, y, z = ct2.ic_plummer_mix(n, 2000, 3, 0.02, 0.2, 17)
    rung = np.where(rng.random(n) < 0.45, 0, 1).astype(np.int8)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, None, None, None, rung)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(2e-3))
    ct2.set_force_options(soft, gm, ct2.KERNEL_PLUMMER)
    ct2.particle_sort_by_rung(1)
    first, _ = ct2.particle_current_range()
    assert n - first >= 2048 and (n - first) % 1024 != 0
    p0 = ct2.particles_download()
    ct2.kick(None, getattr(ct2, flag_name) | ct2.KICK_STORE)
    fx, fy, fz, pot = ct2.forces_download()
    ax, ay, az, ap = orc.kick(p0.x, p0.y, p0.z, gm, soft, orc.KERNEL_PLUMMER, sinks=np.arange(first, n))
    W = np.stack([ax, ay, az])
    e = np.linalg.norm(np.stack([fx, fy, fz])[:, first:] - W, axis=0) / np.linalg.norm(W, axis=0)
    assert np.median(e) < 2e-6 and np.percentile(e, 99) < 5e-5, (np.median(e), e.max())
    assert np.max(np.abs(pot[first:] / ap - 1)) < 1e-6
    p1 = ct2.particles_download()
    assert np.array_equal(p1.vx[:first], p0.vx[:first])


class Trace:
    def __init__(self, lines):
        self.lines, self.i = lines, 0

    def expect(self, prefix):
        line = self.lines[self.i]
        self.i += 1
        assert line.startswith(prefix), (self.i, line, prefix)
        return line


def replay_kick_step(tr, minrung, max_rung, first, t0, minrung0, nocrop):
    """kick_step's control flow (main.cpp:8-92) restated: which levels are
    kicked with which flags, where the rung range is pushed, sorted and
    popped, when the top level is clipped, and the drifts."""
    levels = list(range(max_rung, minrung, -1)) + [minrung]
    ascending, clip_top = True, False
    li = 0
    while li < len(levels):
        level = levels[li]
        top = level == minrung
        if top:
            ascending = False
        if not ascending and not top:
            tr.expect("push")
        if not ascending:
            tr.expect(f"sort_by_rung {level}")
        if top and minrung == minrung0:
            counts = [int(v) for v in tr.expect("rung_counts").split()[1:]]
            if len(counts) > minrung0 + 1 and 3 * sum(counts[minrung0 + 1:]) > counts[minrung0]:
                clip_top = not nocrop
        clipped = clip_top and top
        tr.expect("to_grid ")  # main.cpp:49: the cells of the range are rebuilt before every kick
        f = tr.expect(f"kick rung {level} ").split()
        got = {f[k]: int(f[k + 1]) for k in (3, 5, 7, 9)}
        assert got == {"asc": int(ascending or top), "desc": int(False if clipped else (not ascending or top)),
                       "top": int((clip_top and level == minrung0 + 1) or top), "first": int(first)}, (f, level)
        reached = int(f[-1])
        if clipped:
            tr.expect(f"set_minrung {minrung0 + 1}")
            levels.append(minrung0 + 1)
        elif (not ascending or top) and reached > level:
            levels.append(level + 1)
        if (not ascending or top) and not clipped:
            dt = np.float32(tr.expect("drift").split()[1])  # 10 digits round-trip FP32
            assert dt == np.float32(t0 / 2 ** level)
        if ascending and not top:
            tr.expect("pop")
        li += 1
    return minrung + 1 if clip_top else minrung


def test_kick_step_schedule():
    r = subprocess.run([EXE, "rungs"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    tr = Trace(lines)
    minrung = 0
    for step in range(5):
        f = tr.expect(f"step {step} ").split()
        assert int(f[3]) == minrung
        minrung = replay_kick_step(tr, minrung, int(f[5]), step == 0, 0.01, 0, False)
    f = tr.expect("end").split()
    assert int(f[2]) == minrung and tr.i == len(lines)
    text = r.stdout
    assert "push" in text and "pop" in text and int(f[4]) >= 3  # several rungs were exercised
