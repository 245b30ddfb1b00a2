"""Python face of libct2hip.so, named after the reference's free functions.

The reference keeps its state in process globals (particle.hpp:22-28,
options.cpp:116) and exposes free functions; this module does the same over
the C ABI: ``particle_allocate``, ``kick``, ``drift``, ``fft3d`` ...  Arrays
cross as numpy; every call checks the C status and raises ``Ct2Error``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import Ct2Error, ForceOptions, KickParams  # noqa: F401  (re-exported)

KERNEL_PLUMMER = L.KERNEL_PLUMMER
KERNEL_CT2 = L.KERNEL_CT2
KICK_NAIVE = L.KICK_NAIVE
KICK_STORE = L.KICK_STORE
KICK_ORDERED = L.KICK_ORDERED
KICK_SYMMETRIC = L.KICK_SYMMETRIC
PATH_DIRECT, PATH_PM, PATH_PM_NEAR, PATH_PM_MLP = L.PATH_DIRECT, L.PATH_PM, L.PATH_PM_NEAR, L.PATH_PM_MLP
PATH_MM = L.PATH_MM


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags.c_contiguous, "arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def init(device: int = 0) -> None:
    L.check(L.load().ct2_init(device))


def finalize() -> None:
    L.check(L.load().ct2_finalize())


def synchronize() -> None:
    L.check(L.load().ct2_synchronize())


def stream() -> int:
    s = C.c_void_p()
    L.check(L.load().ct2_get_stream(C.byref(s)))
    return s.value or 0


def set_force_options(hsoft: float, GM: float, kernel: int = KERNEL_PLUMMER) -> None:
    """options.hsoft / options.GM (options.cpp:155,190) and the pair kernel."""
    o = ForceOptions(hsoft, GM, kernel)
    L.check(L.load().ct2_set_force_options(C.byref(o)))


# ---------------------------------------------------------------- particles
def particle_allocate(count: int) -> None:
    """particle_allocate (particle.cpp:273-288) on the device."""
    L.check(L.load().ct2_particles_alloc(count))


def particle_count() -> int:
    return int(L.load().ct2_particle_count())


def pinned_empty(n: int, dtype) -> np.ndarray:
    """An uninitialised numpy array of n elements in page-locked host memory
    (ct2_host_alloc), freed with ct2_host_free when the array is collected."""
    import weakref

    dt = np.dtype(dtype)
    nbytes = int(n) * dt.itemsize
    ptr = C.c_void_p()
    L.check(L.load().ct2_host_alloc(nbytes, C.byref(ptr)))
    if not ptr.value:
        return np.empty(n, dt)
    buf = (C.c_char * nbytes).from_address(ptr.value)
    a = np.frombuffer(buf, dtype=dt, count=n)
    weakref.finalize(buf, L.load().ct2_host_free, ptr.value)
    return a


def particles_upload(x=None, y=None, z=None, vx=None, vy=None, vz=None, rung=None) -> None:
    n = particle_count()
    arrs = []
    for a, dt in ((x, np.uint32), (y, np.uint32), (z, np.uint32), (vx, np.float32),
                  (vy, np.float32), (vz, np.float32), (rung, np.int8)):
        if a is None:
            arrs.append(None)
            continue
        a = np.ascontiguousarray(a, dtype=dt)
        if a.size != n:
            raise ValueError(f"array of {a.size} elements for {n} particles")
        arrs.append(a)
    L.check(L.load().ct2_particles_upload(*[_p(a) for a in arrs]))


@dataclass
class Particles:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    rung: np.ndarray


def particles_download() -> Particles:
    n = particle_count()
    out = [np.empty(n, np.uint32) for _ in range(3)] + [np.empty(n, np.float32) for _ in range(3)]
    out.append(np.empty(n, np.int8))
    L.check(L.load().ct2_particles_download(*[_p(a) for a in out]))
    return Particles(*out)


# ---------------------------------------------------------------- kick / drift
def kick_params(ascending=False, descending=False, top=False, first=False, rung=0, t0=1.0, eta=0.15,
                max_dt=1.0, scale=1.0) -> KickParams:
    return KickParams(int(ascending), int(descending), int(top), int(first), rung, t0, eta, max_dt, scale)


def kick(params: KickParams | None = None, flags: int = 0) -> int:
    """int kick(const kick_params&) (kick.cpp:105-125); returns the max rung.
    With ``params`` None the kick only evaluates forces."""
    if params is None:
        params = kick_params()
    mr = C.c_int(0)
    L.check(L.load().ct2_kick(C.byref(params), flags, C.byref(mr)))
    return mr.value


def kick_noalloc(params: KickParams | None = None, flags: int = 0) -> int:
    """ct2_kick_noalloc: the ordered kick with no device allocation (at most
    NOALLOC_MAX_SINKS active sinks); returns the max rung."""
    if params is None:
        params = kick_params()
    mr = C.c_int(0)
    L.check(L.load().ct2_kick_noalloc(C.byref(params), flags, C.byref(mr)))
    return mr.value


NOALLOC_MAX_SINKS = 1 << 18


def forces_download(out=None):
    """The stored force and potential, into `out` (four float32 arrays of
    the particle count, e.g. from pinned_empty) or new arrays."""
    n = particle_count()
    if out is None:
        out = [np.empty(n, np.float32) for _ in range(4)]
    elif len(out) != 4 or any(a.dtype != np.float32 or a.size != n or not a.flags.c_contiguous for a in out):
        raise ValueError("forces_download needs four contiguous float32 arrays of the particle count")
    L.check(L.load().ct2_forces_download(*[_p(a) for a in out]))
    return tuple(out)


def kick_sample_fp64(idx, periodic: bool = False):
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = [np.empty(idx.size, np.float64) for _ in range(4)]
    L.check(L.load().ct2_kick_sample_fp64(_p(idx), idx.size, int(periodic), *[_p(a) for a in out]))
    return tuple(out)


def drift(dt: float, scale: float) -> None:
    """void drift(float dt, float scale) (drift.cpp:4-18)."""
    L.check(L.load().ct2_drift(dt, scale))


# ---------------------------------------------------------------- rungs
def particle_sort_by_rung(minrung: int) -> None:
    """particle_sort_by_rung (particle.cpp:320-340), stable."""
    L.check(L.load().ct2_particle_sort_by_rung(minrung))


def particle_push_rungs() -> None:
    L.check(L.load().ct2_particle_push_rungs())


def particle_pop_rungs() -> None:
    L.check(L.load().ct2_particle_pop_rungs())


def particle_current_range() -> tuple[int, int]:
    a, b = C.c_size_t(0), C.c_size_t(0)
    L.check(L.load().ct2_particle_current_range(C.byref(a), C.byref(b)))
    return a.value, b.value


def particle_rung_counts() -> list[int]:
    counts = np.zeros(128, np.uint64)
    nr = C.c_int(0)
    L.check(L.load().ct2_particle_rung_counts(_p(counts), 128, C.byref(nr)))
    return [int(v) for v in counts[:nr.value]]


def particle_set_minrung(minrung: int) -> None:
    L.check(L.load().ct2_particle_set_minrung(minrung))


def particles_to_grid(ncell: int) -> np.ndarray:
    """Sorts the current rung range by cell (ct2_particles_to_grid); returns
    the ncell^3 + 1 cell starts (container indices)."""
    starts = np.zeros(ncell ** 3 + 1, np.uint64)
    L.check(L.load().ct2_particles_to_grid(ncell, _p(starts)))
    return starts


def debug_fail_allocations(after: int, count: int) -> None:
    """After `after` more device allocations succeed, the next `count` fail
    with CT2_ERR_NOMEM (test hook; count 0 turns it off)."""
    L.check(L.load().ct2_debug_fail_allocations(after, count))


# ---------------------------------------------------------------- particle mesh
_pm_ng = 0


def pm_init(ng: int, rs_cells: float = 1.25) -> None:
    """Allocates the PM grids and the FP64 Green's table (ct2_pm_init)."""
    global _pm_ng
    L.check(L.load().ct2_pm_init(ng, rs_cells))
    _pm_ng = ng


def pm_near(rcut_rs: float) -> None:
    """Adds the near-field direct sum (pairs closer than rcut_rs * rs) to every
    PM force evaluation; 0 switches it off (ct2_pm_near)."""
    L.check(L.load().ct2_pm_near(rcut_rs))


def pm_assignment(order: int) -> None:
    """Mass assignment of the PM force: 1 CIC (default), 2 TSC (ct2_pm_assignment)."""
    L.check(L.load().ct2_pm_assignment(order))


def pm_near_cells() -> int:
    return int(L.load().ct2_pm_near_cells())


def near_kernel_table(rs_u: float, rcut_u: float, nt: int = 64):
    """The cubic pieces of the near-field kernel's filtered part (host
    function): (k4, p4), each float32 of shape (nt, 4)."""
    k4, p4 = np.zeros((nt, 4), np.float32), np.zeros((nt, 4), np.float32)
    L.check(L.load().ct2_near_kernel_table(rs_u, rcut_u, nt, k4.ctypes.data_as(C.c_void_p),
                                           p4.ctypes.data_as(C.c_void_p)))
    return k4, p4


def pm_forces():
    """PM force and potential of every particle: (fx, fy, fz, pot)."""
    L.check(L.load().ct2_pm_forces())
    return forces_download()


def kick_pm(params: KickParams | None = None, flags: int = 0) -> int:
    if params is None:
        params = kick_params()
    mr = C.c_int(0)
    L.check(L.load().ct2_kick_pm(C.byref(params), flags, C.byref(mr)))
    return mr.value


# ---------------------------------------------------------------- multipole mesh
_mm = {"nc": 0, "order": 0}


def mm_ncomp(order: int) -> int:
    return (order + 1) * (order + 2) * (order + 3) // 6


def mm_cells_for(n: int, per_cell: int = 8) -> int:
    """The power-of-two cell grid with about `per_cell` (at least) particles per
    cell, at least 8 cells a side (ct2_mm_cells_for); at 8 per cell the
    near-set pairs and the far field cost about the same."""
    nc = C.c_int(0)
    L.check(L.load().ct2_mm_cells_for(n, per_cell, C.byref(nc)))
    return nc.value


def mm_init(nc: int, order: int) -> None:
    """ct2's multipole-mesh far field on nc^3 cells with Cartesian order `order`."""
    L.check(L.load().ct2_mm_init(nc, order))
    _mm.update(nc=nc, order=order)


def mm_forces():
    """Multipole mesh + near-set direct pairs for every particle: (fx, fy, fz, pot)."""
    L.check(L.load().ct2_mm_forces())
    return forces_download()


def kick_mm(params: KickParams | None = None, flags: int = 0) -> int:
    if params is None:
        params = kick_params()
    mr = C.c_int(0)
    L.check(L.load().ct2_kick_mm(C.byref(params), flags, C.byref(mr)))
    return mr.value


def mm_table_download() -> np.ndarray:
    nc, k = _mm["nc"], mm_ncomp(_mm["order"])
    t = np.empty(k * nc ** 3)
    L.check(L.load().ct2_mm_table_download(_p(t)))
    return t.reshape(k, nc, nc, nc)


def pm_density_download() -> np.ndarray:
    g = np.empty(_pm_ng ** 3, np.uint64)
    L.check(L.load().ct2_pm_density_download(_p(g)))
    return g.reshape((_pm_ng,) * 3)


def pm_grid_download(which: int) -> np.ndarray:
    g = np.empty(_pm_ng ** 3, np.complex64)
    L.check(L.load().ct2_pm_grid_download(which, _p(g)))
    return g.reshape((_pm_ng,) * 3)


# ---------------------------------------------------------------- learned correction
MLP_KIN = 128


def mlp_num_params() -> int:
    return int(L.load().ct2_mlp_num_params())


def mlp_init(seed: int = 1) -> None:
    L.check(L.load().ct2_mlp_init(seed))


def mlp_get_weights() -> np.ndarray:
    w = np.empty(mlp_num_params(), np.float32)
    L.check(L.load().ct2_mlp_get_weights(_p(w), w.size))
    return w


def mlp_set_weights(w) -> None:
    w = np.ascontiguousarray(w, dtype=np.float32)
    L.check(L.load().ct2_mlp_set_weights(_p(w), w.size))


def mlp_set_mesh(rs_cells: float, order: int, rcut_rs: float = 0.0, particles_per_node: float = 0.0) -> None:
    """Records the mesh the current weights were trained for (ct2_mlp_set_mesh):
    ct2_kick_pm_mlp falls back to the direct sum on any other mesh."""
    L.check(L.load().ct2_mlp_set_mesh(rs_cells, order, rcut_rs, particles_per_node))


def mlp_mesh_of(meta: dict) -> tuple[float, int, float, float]:
    """(rs_cells, order, rcut_rs, particles per node) of a trained model's
    metadata (train_mlp.py)."""
    ppn = (meta["side"] / meta["ng"]) ** 3 if meta.get("side") else 0.0
    return float(meta["rs_cells"]), int(meta.get("order", 1)), float(meta.get("near_rcut_rs", 0.0)), float(ppn)


def last_force_path() -> int:
    """CT2_PATH_* of the last kick (ct2_last_force_path)."""
    return int(L.load().ct2_last_force_path())


def mlp_forward(X) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.empty((X.shape[0], 3), np.float32)
    L.check(L.load().ct2_mlp_forward(_p(X), X.shape[0], _p(Y)))
    return Y


def mlp_grad(X, T):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.empty(mlp_num_params(), np.float32)
    loss = C.c_float()
    L.check(L.load().ct2_mlp_grad(_p(X), _p(T), X.shape[0], _p(g), C.byref(loss)))
    return loss.value, g


def mlp_features(idx) -> np.ndarray:
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    X = np.empty((idx.size, MLP_KIN), np.float32)
    L.check(L.load().ct2_mlp_features(_p(idx), idx.size, _p(X)))
    return X


def mlp_data_reset(capacity: int) -> None:
    L.check(L.load().ct2_mlp_data_reset(capacity))


def mlp_data_size() -> int:
    return int(L.load().ct2_mlp_data_size())


def mlp_data_append(idx) -> None:
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    for a in range(0, idx.size, 1 << 20):
        part = np.ascontiguousarray(idx[a:a + (1 << 20)])
        L.check(L.load().ct2_mlp_data_append(_p(part), part.size))


def mlp_train(steps: int, batch: int = 1024, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
    losses = np.empty(steps, np.float32)
    L.check(L.load().ct2_mlp_train(steps, batch, lr, seed, _p(losses)))
    return losses


def pm_mlp_forces():
    """PM force plus the learned correction: (fx, fy, fz, pot)."""
    L.check(L.load().ct2_pm_mlp_forces())
    return forces_download()


def kick_pm_mlp(params: KickParams | None = None, flags: int = 0) -> int:
    if params is None:
        params = kick_params()
    mr = C.c_int(0)
    L.check(L.load().ct2_kick_pm_mlp(C.byref(params), flags, C.byref(mr)))
    return mr.value


# ---------------------------------------------------------------- device buffers / FFT
class DeviceBuffer:
    """A device allocation made through the library (freed on close/del)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        L.check(L.load().ct2_dev_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        L.check(L.load().ct2_memcpy_h2d(self.ptr, _p(a), a.nbytes))

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        L.check(L.load().ct2_memcpy_d2h(_p(a), self.ptr, a.nbytes))
        return a

    def close(self) -> None:
        if self.ptr:
            L.check(L.load().ct2_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr:
                L.load().ct2_dev_free(self.ptr)
        except Exception:
            pass


def fft3d_device(buf: DeviceBuffer, n: int, dir: int = 1, f64: bool = False) -> None:
    fn = L.load().ct2_fft3d_c2c_f64 if f64 else L.load().ct2_fft3d_c2c
    L.check(fn(buf.ptr, n, dir))


def fft3d(a: np.ndarray, dir: int = 1) -> np.ndarray:
    """3-D complex FFT of a host cube through the device (complex64 -> FP32
    kernels, complex128 -> FP64 kernels); unnormalised, FFTW sign."""
    n = a.shape[0]
    assert a.shape == (n, n, n) and a.dtype in (np.complex64, np.complex128)
    buf = DeviceBuffer(a.nbytes)
    try:
        buf.upload(a)
        fft3d_device(buf, n, dir, a.dtype == np.complex128)
        return buf.download(np.empty_like(a))
    finally:
        buf.close()


# ---------------------------------------------------------------- ct2's FFT API on the device
def _through_device(arrays, call):
    """Uploads float64 host arrays, runs call(*pointers), synchronises and
    returns the downloaded arrays."""
    bufs = [DeviceBuffer(max(8, a.nbytes)) for a in arrays]
    try:
        for b, a in zip(bufs, arrays):
            b.upload(np.ascontiguousarray(a, np.float64))
        L.check(call(*[b.ptr for b in bufs]))
        synchronize()
        return [b.download(np.empty_like(np.ascontiguousarray(a, np.float64))) for b, a in zip(bufs, arrays)]
    finally:
        for b in bufs:
            b.close()


def fft_complex(X: np.ndarray, N: int, dir: int = 1, lanes: int = 1) -> np.ndarray:
    """fft_complex(complex<simd_f64>*, N, dir) (fft_complex_vector.cpp:9-65) on
    the GPU: X holds batches of 2*lanes*N doubles, bit-reversed input."""
    batch = X.size // (2 * lanes * N)
    return _through_device([X], lambda p: L.load().ct2_fft_complex_1d(p, N, dir, lanes, batch))[0]


def fft_real(X: np.ndarray, N: int, dir: int = 1, lanes: int = 1) -> np.ndarray:
    """fft_real(double*/simd_f64*, N, dir) (fft_real_*.cpp) on the GPU:
    bit-reversed real input -> half-complex output, batches of lanes*N."""
    batch = X.size // (lanes * N)
    return _through_device([X], lambda p: L.load().ct2_fft_real_1d(p, N, dir, lanes, batch))[0]


def scramble(X: np.ndarray, N1: int, N2: int, N3: int) -> np.ndarray:
    return _through_device([X], lambda p: L.load().ct2_scramble(p, N1, N2, N3))[0]


def transpose_yz(X: np.ndarray, N: int) -> np.ndarray:
    return _through_device([X], lambda p: L.load().ct2_transpose_yz(p, N))[0]


def transpose_r2c(X: np.ndarray, N: int) -> np.ndarray:
    out = np.zeros(2 * N * N * (N // 2 + 1))
    return _through_device([X, out], lambda a, b: L.load().ct2_transpose_r2c(a, b, N))[1]


def transpose_c2r(Y: np.ndarray, N: int) -> np.ndarray:
    out = np.zeros(N ** 3)
    return _through_device([Y, out], lambda a, b: L.load().ct2_transpose_c2r(a, b, N))[1]


def fft_r2c_3d(X: np.ndarray) -> np.ndarray:
    """fft_r2c_3d(in, out, N): real N^3 -> planar half spectrum (2 N^2 (N/2+1))."""
    N = round(X.size ** (1 / 3))
    out = np.zeros(2 * N * N * (N // 2 + 1))
    return _through_device([X, out], lambda a, b: L.load().ct2_fft_r2c_3d(a, b, N))[1]


def fft_c2r_3d(Y: np.ndarray, N: int) -> np.ndarray:
    """fft_c2r_3d(in, out, N): planar half spectrum -> real N^3, unnormalised."""
    out = np.zeros(N ** 3)
    return _through_device([Y, out], lambda a, b: L.load().ct2_fft_c2r_3d(a, b, N))[1]


def twiddles(N: int):
    """cos_twiddles, sin_twiddles, inv_sin_twiddles (twiddles.cpp:6-54)."""
    out = [np.empty(N) for _ in range(3)]
    L.check(L.load().ct2_twiddles(N, *[_p(a) for a in out]))
    return tuple(out)


# ---------------------------------------------------------------- initial conditions
def ic_lattice(n: int, sigma_cells: float, seed: int):
    N = n ** 3
    out = [np.empty(N, np.uint32) for _ in range(3)]
    L.check(L.load().ct2_ic_lattice(n, sigma_cells, seed, *[_p(a) for a in out]))
    return tuple(out)


def ic_plummer_mix(n: int, n_sph: int, nspheres: int, a: float, rmax: float, seed: int):
    out = [np.empty(n, np.uint32) for _ in range(3)]
    L.check(L.load().ct2_ic_plummer_mix(n, n_sph, nspheres, a, rmax, seed, *[_p(x) for x in out]))
    return tuple(out)


def ic_grf(n: int, rms_cells: float, spectral_index: float = -2.0, seed: int = 1, vfac: float = 0.0) -> None:
    """Zel'dovich lattice from a Gaussian random field, written into the
    device container (which must hold n^3 particles)."""
    L.check(L.load().ct2_ic_grf(n, rms_cells, spectral_index, seed, vfac))


# ---------------------------------------------------------------- cosmology
Cosmology = L.Cosmology


def cosmology(H: float, omega_m: float = 1.0, omega_r: float = 0.0, omega_lam: float = 0.0):
    return L.Cosmology(H, omega_m, omega_r, omega_lam)


def cosmos_update(c, adotdot: float, adot: float, a: float, dt0: float):
    """cosmos_update (cosmology.cpp:38-44): returns (adotdot, adot, a)."""
    v = [C.c_double(adotdot), C.c_double(adot), C.c_double(a)]
    L.check(L.load().ct2_cosmos_update(C.byref(c), C.byref(v[0]), C.byref(v[1]), C.byref(v[2]), dt0))
    return v[0].value, v[1].value, v[2].value


def cosmos_dadt(c, a: float) -> float:
    return L.load().ct2_cosmos_dadt(C.byref(c), a)


def cosmos_conformal_time(c, a0: float, a1: float) -> float:
    return L.load().ct2_cosmos_conformal_time(C.byref(c), a0, a1)


def cosmos_time(c, a0: float, a1: float) -> float:
    return L.load().ct2_cosmos_time(C.byref(c), a0, a1)


def cosmos_kick_factor(c, a0: float, a1: float) -> float:
    return L.load().ct2_cosmos_kick_factor(C.byref(c), a0, a1)


def cosmos_drift_factor(c, a0: float, a1: float) -> float:
    return L.load().ct2_cosmos_drift_factor(C.byref(c), a0, a1)


# ---------------------------------------------------------------- options, code units
Options = L.Options


def options(**kw) -> "L.Options":
    """The options block with the reference's defaults (options.cpp:51-61),
    overridden by ``kw``, and its derived fields (options.cpp:88-95) filled."""
    o = L.Options()
    L.check(L.load().ct2_options_defaults(C.byref(o)))
    for k, v in kw.items():
        if k not in {f[0] for f in L.Options._fields_}:
            raise KeyError(f"unknown option {k!r}")
        setattr(o, k, v)
    L.check(L.load().ct2_options_derive(C.byref(o)))
    return o


def options_cosmology(o) -> "L.Cosmology":
    """The cosmology block of an options block: H = H0 code_to_s hubble in code
    time units (cosmology.cpp:15)."""
    c = L.Cosmology()
    L.check(L.load().ct2_options_cosmology(C.byref(o), C.byref(c)))
    return c
