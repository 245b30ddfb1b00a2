"""The reference-side binding: ct2's own kick()/drift() signatures and
particle globals (ct2_amd/compat) driven from C++ the way kick_step drives
them (main.cpp:62-81), checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_bin", "compat_driver")


def lcg_positions(n, seed=12345):
    s = seed
    out = np.empty((n, 3), np.uint32)
    for i in range(n):
        for d in range(3):
            s = (s * 1664525 + 1013904223) % 2 ** 32
            out[i, d] = s
    return out.T


@pytest.mark.parametrize("kernel", [None, 0])
def test_kick_and_drift_through_ct2_signatures(orc, kernel):
    """Without a kernel argument ct2_gpu_set_options selects ct2's own
    uniform-density-sphere kernel (kick.cpp:52-60); 0 selects Plummer."""
    n = 512
    args = [EXE, str(n)] + ([] if kernel is None else [str(kernel)])
    r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    max_rung = int(lines[0].split()[1])
    rows = [l.split() for l in lines[1:]]
    x, y, z = (np.array([int(v[a]) for v in rows], np.uint32) for a in range(3))
    vx, vy, vz = (np.array([float(v[a]) for v in rows], np.float32) for a in range(3, 6))
    rung = np.array([int(v[6]) for v in rows])
    x0, y0, z0 = lcg_positions(n)
    gm, soft = float(np.float32(1.0 / n)), float(np.float32(0.01))
    ax, ay, az, _ = orc.kick(x0, y0, z0, gm, soft, orc.KERNEL_CT2 if kernel is None else orc.KERNEL_PLUMMER)
    # ascending (not first) and descending both half-kick: v = 2 F dt/2
    dtk = float(np.float32(0.5) * np.float32(0.01))
    V = np.stack([vx, vy, vz]).astype(np.float64)
    W = 2 * dtk * np.stack([ax, ay, az])
    err = np.linalg.norm(V - W, axis=0) / np.linalg.norm(W, axis=0)  # FP32 force vs FP64, per particle
    assert np.median(err) < 2e-6 and err.max() < 1e-4, (np.median(err), err.max())
    # the drift is bit-exact given the velocities the kick produced
    wx, wy, wz = orc.drift(x0, y0, z0, vx, vy, vz, 0.01, 1.0)
    assert np.array_equal(x, wx) and np.array_equal(y, wy) and np.array_equal(z, wz)
    assert max_rung == rung.max()


def test_fft_api_through_ct2_signatures(orc):
    """scramble + fft_real and fft_r2c_3d / fft_c2r_3d through ct2's own
    signatures (fft.hpp:18-29), host arrays in and out, as fft_test.cpp calls
    them; checked against the oracle and the round trip."""
    r = subprocess.run([EXE, "fft"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = np.array([float(t) for t in r.stdout.split()])
    N, M = 1024, 16
    h = M // 2 + 1
    s = 777
    vals = []
    for _ in range(N + M ** 3):
        s = (s * 1664525 + 1013904223) % 2 ** 32
        vals.append(s / 2 ** 32 - 0.5)
    x, cube = np.array(vals[:N]), np.array(vals[N:]).reshape(M, M, M)
    want = orc.fft_real(orc.scramble(x, 1, N, 1), N)
    got, v = v[:N], v[N:]
    assert np.linalg.norm(got - want) < 1e-14 * np.linalg.norm(want)
    spec, v = v[:2 * M * M * h], v[2 * M * M * h:]
    ws = orc.fft_r2c_3d(cube)
    assert np.linalg.norm(spec - ws) < 1e-14 * np.linalg.norm(ws)
    back, v = v[:M ** 3], v[M ** 3:]
    assert np.allclose(back / M ** 3, cube.ravel(), atol=1e-14)
    assert np.allclose(v, orc.twiddles(8)[0], atol=1e-16)


def test_learned_correction_through_kick(ct2, tmp_path):
    """kick() switched to PM + the learned correction (ct2_gpu_use_learned,
    TSC 32^3, split 1 cell) gives, bit for bit, the velocities and rungs of
    the library's own ct2_kick_pm_mlp on the same particles and weights."""
    nw = ct2.mlp_num_params()
    w = (np.random.default_rng(5).standard_normal(nw) * 0.05).astype(np.float32)
    path = tmp_path / "weights.f32"
    w.tofile(path)
    r = subprocess.run([EXE, "learned", str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    max_rung = int(lines[0].split()[1])
    rows = [l.split() for l in lines[1:]]
    got = np.array([[float(v) for v in row[:3]] for row in rows], np.float32).T
    got_rung = np.array([int(row[3]) for row in rows])
    n = got.shape[1]
    x, y, z = lcg_positions(n, 777)
    zf = np.zeros(n, np.float32)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    ct2.set_force_options(float(np.float32(0.01)), float(np.float32(1.0 / n)), 1)  # the binding's default kernel
    try:
        ct2.pm_assignment(2)
        ct2.pm_init(32, 1.0)
        ct2.pm_near(0.0)
        ct2.mlp_init(1)
        ct2.mlp_set_weights(w)
        params = ct2.kick_params(ascending=True, descending=True, top=True, first=False, rung=0, t0=0.01, eta=0.15,
                                 max_dt=0.01, scale=1.0)
        want_max = ct2.kick_pm_mlp(params)
        p = ct2.particles_download()
    finally:
        ct2.pm_assignment(1)
    assert max_rung == want_max
    assert np.array_equal(got, np.stack([p.vx, p.vy, p.vz]))
    assert np.array_equal(got_rung, p.rung.astype(int))
    assert np.abs(got).max() > 0


def test_grid_accessors_through_ct2_signatures(ct2):
    """particles_init, particle_sort_by_rung, particle_to_grid and the cell
    accessors of particle.hpp:58-70 through the binding: the rung range is
    sorted stably by cell (ct2's Ncell = cbrt(count / ppcell) + 4), each cell's
    range holds exactly its particles, and the neighbour list of a cell has
    one entry per offset with i0^2 + j0^2 + k0^2 <= bwidth^2."""
    r = subprocess.run([EXE, "grid"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    head = lines[0].split()
    first, last, nc, nnb = int(head[1]), int(head[2]), int(head[4]), int(head[6])
    n = 20000
    count = last - first
    assert nc == int((count // 40) ** (1.0 / 3.0) + 4)
    b = 2
    assert nnb == sum(1 for i in range(-b, b + 1) for j in range(-b, b + 1) for k in range(-b, b + 1)
                      if i * i + j * j + k * k <= b * b)
    cells = np.array([[int(t) for t in l.split()] for l in lines[1:1 + nc ** 3]])
    rows = [l.split() for l in lines[1 + nc ** 3:]]
    x, y, z = (np.array([int(v[a]) for v in rows], np.uint64) for a in range(3))
    tag = np.array([float(v[3]) for v in rows]).astype(np.int64)
    rung = np.array([int(v[4]) for v in rows])
    # particles_init: the library's seeded uniform set
    x0, y0, z0 = ct2.ic_plummer_mix(n, 0, 0, 0.0, 0.0, 1)
    assert np.array_equal(x, x0[tag]) and np.array_equal(y, y0[tag]) and np.array_equal(z, z0[tag])
    r0 = (np.arange(n) // 7) % 2
    assert first == int((r0 == 0).sum()) and last == n
    # the slow particles keep the stable partition's order, the range is in cell order
    assert np.array_equal(tag[:first], np.flatnonzero(r0 == 0)) and (rung[first:] == 1).all()
    key = ((x[first:] * nc >> 32) * nc + (y[first:] * nc >> 32)) * nc + (z[first:] * nc >> 32)
    assert np.all(np.diff(key.astype(np.int64)) >= 0)
    part = np.flatnonzero(r0 == 1)
    kk0 = ((x0[part].astype(np.uint64) * nc >> 32) * nc + (y0[part].astype(np.uint64) * nc >> 32)) * nc + \
        (z0[part].astype(np.uint64) * nc >> 32)
    assert np.array_equal(tag[first:], part[np.argsort(kk0, kind="stable")])
    starts = first + np.searchsorted(key, np.arange(nc ** 3), side="left")
    ends = first + np.searchsorted(key, np.arange(nc ** 3), side="right")
    assert np.array_equal(cells[:, 0], starts) and np.array_equal(cells[:, 1], ends)


def test_multipole_mesh_through_kick(ct2):
    """kick() switched to ct2's own structure (ct2_gpu_use_multipoles: direct
    pairs over the cells with |offset|^2 <= 10 + the multipole mesh, 16^3
    cells, order 6) gives, bit for bit, the velocities and rungs of the
    library's own ct2_kick_mm on the same particles."""
    r = subprocess.run([EXE, "mm"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    max_rung = int(lines[0].split()[1])
    rows = [l.split() for l in lines[1:]]
    got = np.array([[float(v) for v in row[:3]] for row in rows], np.float32).T
    got_rung = np.array([int(row[3]) for row in rows])
    n = got.shape[1]
    x, y, z = lcg_positions(n, 4242)
    zf = np.zeros(n, np.float32)
    ct2.particle_allocate(n)
    ct2.particles_upload(x, y, z, zf, zf, zf, np.zeros(n, np.int8))
    ct2.set_force_options(float(np.float32(0.01)), float(np.float32(1.0 / n)), 1)  # the binding's default kernel
    ct2.mm_init(16, 6)
    params = ct2.kick_params(ascending=True, descending=True, top=True, first=False, rung=0, t0=0.01, eta=0.15,
                             max_dt=0.01, scale=1.0)
    want_max = ct2.kick_mm(params)
    p = ct2.particles_download()
    assert max_rung == want_max
    assert np.array_equal(got, np.stack([p.vx, p.vy, p.vz]))
    assert np.array_equal(got_rung, p.rung.astype(int))
    assert np.abs(got).max() > 0
