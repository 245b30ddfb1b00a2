"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path in ``ct2_amd``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module; the product never does.

Each wrapper names the reference file:line its C function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

KERNEL_PLUMMER = 0
KERNEL_CT2 = 1


def build() -> None:
    """Compile liboracle.so from the C sources in this folder."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class KickParams(C.Structure):
    """kick_params, kick.hpp:10-20."""

    _fields_ = [("ascending", C.c_int), ("descending", C.c_int), ("top", C.c_int),
                ("first", C.c_int), ("rung", C.c_int), ("t0", C.c_float), ("eta", C.c_float),
                ("max_dt", C.c_float), ("scale", C.c_float)]


class Cosmo(C.Structure):
    _fields_ = [("H", C.c_double), ("omega_m", C.c_double), ("omega_r", C.c_double),
                ("omega_lam", C.c_double)]


_P = C.c_void_p


def _declare(L):
    sz, i, d, f = C.c_size_t, C.c_int, C.c_double, C.c_float
    L.orc_kick.argtypes = [sz, _P, _P, _P, sz, _P, d, d, i, _P, _P, _P, _P]
    L.orc_kick_ld.argtypes = L.orc_kick.argtypes
    L.orc_kick_epilogue.argtypes = [sz, _P, _P, _P, _P, _P, _P, _P, C.POINTER(KickParams), f]
    L.orc_kick_epilogue.restype = i
    L.orc_fixed32_from_float.argtypes = [f]
    L.orc_fixed32_from_float.restype = C.c_uint32
    L.orc_drift.argtypes = [sz, _P, _P, _P, _P, _P, _P, f, f]
    L.orc_scramble.argtypes = [_P, i, i, i]
    L.orc_twiddles.argtypes = [i, _P, _P, _P]
    L.orc_transpose_yz.argtypes = [_P, i]
    L.orc_transpose_r2c.argtypes = [_P, _P, i]
    L.orc_transpose_c2r.argtypes = [_P, _P, i]
    L.orc_fft_complex.argtypes = [_P, i, i, i]
    L.orc_fft_real.argtypes = [_P, i, i, i]
    L.orc_fft3d.argtypes = [_P, i, i]
    L.orc_fft_r2c_3d.argtypes = [_P, _P, i]
    L.orc_fft_c2r_3d.argtypes = [_P, _P, i]
    L.orc_cic_deposit.argtypes = [sz, _P, _P, _P, i, _P]
    L.orc_mesh_deposit.argtypes = [sz, _P, _P, _P, i, i, _P]
    L.orc_pm_green_table.argtypes = [i, d, d, i, _P]
    L.orc_pm_forces.argtypes = [sz, _P, _P, _P, i, d, d, i, _P, _P, _P, _P]
    L.orc_ewald_kick.argtypes = [sz, _P, _P, _P, sz, _P, d, d, d, i, i, _P, _P, _P, _P]
    L.orc_pm_near.argtypes = [sz, _P, _P, _P, sz, _P, d, d, i, d, d, _P, _P, _P, _P]
    L.orc_mm_ncomp.argtypes = [i]
    L.orc_mm_ncomp.restype = i
    L.orc_mm_index.argtypes = [i, _P]
    L.orc_mm_green_point.argtypes = [i, i, d, i, _P, _P]
    L.orc_mm_green_table.argtypes = [i, i, d, i, _P]
    L.orc_mm_forces.argtypes = [sz, _P, _P, _P, i, i, d, i, d, d, i, i, _P, _P, _P, _P]
    L.orc_mlp_num_params.restype = sz
    L.orc_bf16_round.argtypes = [d]
    L.orc_bf16_round.restype = d
    L.orc_gelu.argtypes = [sz, _P, _P, _P]
    L.orc_mlp_forward.argtypes = [_P, sz, _P, i, _P]
    L.orc_mlp_grad.argtypes = [_P, sz, _P, _P, i, _P]
    L.orc_mlp_grad.restype = d
    L.orc_set_threads.argtypes = [i]
    L.orc_max_threads.restype = i
    L.orc_cosmos_update.argtypes = [C.POINTER(Cosmo), _P, _P, _P, d]
    L.orc_cosmos_dadt.argtypes = [C.POINTER(Cosmo), d]
    L.orc_cosmos_dadt.restype = d
    L.orc_cosmos_conformal_time.argtypes = [C.POINTER(Cosmo), d, d]
    L.orc_cosmos_conformal_time.restype = d
    L.orc_cosmos_time.argtypes = [C.POINTER(Cosmo), d, d]
    L.orc_cosmos_time.restype = d
    L.orc_options_derive.argtypes = [C.c_uint64, f, f, f, f, _P]


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- kick
def kick(x, y, z, gm, soft, kernel=KERNEL_PLUMMER, sinks=None, long_double=False):
    """FP64 all-pairs kick (kick.cpp:36-74).  Returns (ax, ay, az, pot) for the
    sinks (all particles when ``sinks`` is None)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    n = x.size
    if sinks is None:
        m, sp = n, None
    else:
        sinks = np.ascontiguousarray(sinks, dtype=np.int64)
        m, sp = sinks.size, _p(sinks)
    out = [np.empty(m, np.float64) for _ in range(4)]
    fn = lib().orc_kick_ld if long_double else lib().orc_kick
    fn(n, _p(x), _p(y), _p(z), m, sp, gm, soft, kernel, *[_p(o) for o in out])
    return tuple(out)


def kick_epilogue(fx, fy, fz, vx, vy, vz, rung, params: KickParams, hsoft):
    """Half-kick and rung update (kick.cpp:75-99) in FP32; updates vx, vy, vz
    and rung in place and returns the maximum rung."""
    fx, fy, fz = (np.ascontiguousarray(a, np.float32) for a in (fx, fy, fz))
    for a in (vx, vy, vz):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    assert rung.dtype == np.int8
    return lib().orc_kick_epilogue(fx.size, _p(fx), _p(fy), _p(fz), _p(vx), _p(vy), _p(vz),
                                   _p(rung), C.byref(params), hsoft)


def fixed32_from_float(s: float) -> int:
    return int(lib().orc_fixed32_from_float(s))


def set_threads(t: int) -> None:
    """OpenMP threads of the oracle's parallel loops."""
    lib().orc_set_threads(int(t))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def drift(x, y, z, vx, vy, vz, dt, scale):
    """drift.cpp:4-17 on copies; returns the new (x, y, z)."""
    x, y, z = (np.array(a, dtype=np.uint32, copy=True) for a in (x, y, z))
    vx, vy, vz = (np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz))
    lib().orc_drift(x.size, _p(x), _p(y), _p(z), _p(vx), _p(vy), _p(vz), dt, scale)
    return x, y, z


# ---------------------------------------------------------------- FFT
def scramble(X, N1, N2, N3):
    X = np.array(X, dtype=np.float64, copy=True)
    lib().orc_scramble(_p(X), N1, N2, N3)
    return X


def twiddles(N):
    c, s, i = (np.empty(N) for _ in range(3))
    lib().orc_twiddles(N, _p(c), _p(s), _p(i))
    return c, s, i


def fft_complex(X, N, dir=1, lanes=1):
    """fft_complex_vector.cpp:9-65 on a copy of the interleaved lane layout."""
    X = np.array(X, dtype=np.float64, copy=True)
    lib().orc_fft_complex(_p(X), N, dir, lanes)
    return X


def fft_real(X, N, dir=1, lanes=1):
    """fft_real_scalar.cpp / fft_real_vector.cpp on a copy."""
    X = np.array(X, dtype=np.float64, copy=True)
    lib().orc_fft_real(_p(X), N, dir, lanes)
    return X


def fft3d(Z: np.ndarray, dir=1) -> np.ndarray:
    """3-D complex transform of a complex128 n^3 cube, natural order."""
    Z = np.array(Z, dtype=np.complex128, copy=True)
    n = Z.shape[0]
    lib().orc_fft3d(_p(Z), n, dir)
    return Z


def fft_r2c_3d(X: np.ndarray) -> np.ndarray:
    N = X.shape[0]
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = np.empty(2 * N * N * (N // 2 + 1))
    lib().orc_fft_r2c_3d(_p(X), _p(out), N)
    return out


def fft_c2r_3d(Y: np.ndarray, N: int) -> np.ndarray:
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    out = np.empty(N * N * N)
    lib().orc_fft_c2r_3d(_p(Y), _p(out), N)
    return out.reshape(N, N, N)


def transpose_yz(X, N):
    X = np.array(X, dtype=np.float64, copy=True)
    lib().orc_transpose_yz(_p(X), N)
    return X


def transpose_r2c(X, N):
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = np.empty(2 * N * N * (N // 2 + 1))
    lib().orc_transpose_r2c(_p(X), _p(out), N)
    return out


def transpose_c2r(Y, N):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    out = np.zeros(N * N * N)
    lib().orc_transpose_c2r(_p(Y), _p(out), N)
    return out


# ---------------------------------------------------------------- particle mesh
def cic_deposit(x, y, z, g):
    """CIC mass assignment in 2^-32 particle units (uint64 grid of (2^g)^3)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    ng = 1 << g
    grid = np.empty(ng ** 3, np.uint64)
    lib().orc_cic_deposit(x.size, _p(x), _p(y), _p(z), g, _p(grid))
    return grid.reshape(ng, ng, ng)


def mesh_deposit(x, y, z, g, order=1):
    """Mass assignment of `order` (1 CIC, 2 TSC) in 2^-32 particle units."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    ng = 1 << g
    grid = np.empty(ng ** 3, np.uint64)
    lib().orc_mesh_deposit(x.size, _p(x), _p(y), _p(z), g, order, _p(grid))
    return grid.reshape(ng, ng, ng)


def pm_green_table(ng, gm, rs_cells=1.25, order=1):
    h = ng // 2 + 1
    t = np.empty(h ** 3)
    lib().orc_pm_green_table(ng, gm, rs_cells, order, _p(t))
    return t.reshape(h, h, h)


def pm_forces(x, y, z, g, gm, rs_cells=1.25, order=1):
    """FP64 PM force and potential at the particles: (fx, fy, fz, pot)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    out = [np.empty(x.size) for _ in range(4)]
    lib().orc_pm_forces(x.size, _p(x), _p(y), _p(z), g, gm, rs_cells, order, *[_p(o) for o in out])
    return tuple(out)


def pm_near(x, y, z, gm, soft, rs, rcut, kernel=KERNEL_PLUMMER, sinks=None):
    """FP64 near-field complement of the Gaussian-filtered PM force (pairs
    closer than rcut; rs and rcut in box units): (ax, ay, az, pot)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    if sinks is None:
        m, sp = x.size, None
    else:
        sinks = np.ascontiguousarray(sinks, dtype=np.int64)
        m, sp = sinks.size, _p(sinks)
    out = [np.empty(m) for _ in range(4)]
    lib().orc_pm_near(x.size, _p(x), _p(y), _p(z), m, sp, gm, soft, kernel, rs, rcut, *[_p(o) for o in out])
    return tuple(out)


def ewald_kick(x, y, z, gm, soft, sinks=None, alpha=2.5, nimg=2, kmax=5):
    """FP64 periodic direct sum (Ewald, explicit images): (ax, ay, az, pot)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    if sinks is None:
        m, sp = x.size, None
    else:
        sinks = np.ascontiguousarray(sinks, dtype=np.int64)
        m, sp = sinks.size, _p(sinks)
    out = [np.empty(m) for _ in range(4)]
    lib().orc_ewald_kick(x.size, _p(x), _p(y), _p(z), m, sp, gm, soft, alpha, nimg, kmax, *[_p(o) for o in out])
    return tuple(out)


# ---------------------------------------------------------------- multipole mesh
MM_BETA = 0.5  # Ewald split parameter per cell of the derivative table
MM_NEAR2 = 10  # |cell offset|^2 of the direct set (particle.cpp:136)


def mm_index(P):
    k = lib().orc_mm_ncomp(P)
    out = np.empty(3 * k, np.int32)
    lib().orc_mm_index(P, _p(out))
    return out.reshape(k, 3)


def mm_green_point(P, nc, rho, near=False, beta=MM_BETA):
    k = lib().orc_mm_ncomp(P)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    D = np.empty(k)
    lib().orc_mm_green_point(P, nc, beta, int(near), _p(rho), _p(D))
    return D


def mm_green_table(P, nc, near2=MM_NEAR2, beta=MM_BETA):
    k = lib().orc_mm_ncomp(P)
    Dt = np.empty(k * nc ** 3)
    lib().orc_mm_green_table(P, nc, beta, near2, _p(Dt))
    return Dt.reshape(k, nc, nc, nc)


def mm_forces(x, y, z, nc, P, gm, soft, kernel=KERNEL_PLUMMER, parts=3, near2=MM_NEAR2, beta=MM_BETA):
    """FP64 multipole-mesh force and potential (box units): (fx, fy, fz, pot)."""
    x, y, z = _u32(x), _u32(y), _u32(z)
    out = [np.empty(x.size) for _ in range(4)]
    lib().orc_mm_forces(x.size, _p(x), _p(y), _p(z), nc, P, beta, near2, gm, soft, kernel, parts,
                        *[_p(o) for o in out])
    return tuple(out)


# ---------------------------------------------------------------- learned correction
MLP_KIN, MLP_HID, MLP_KOUT = 128, 256, 32


def mlp_num_params() -> int:
    return int(lib().orc_mlp_num_params())


def bf16_round(v: float) -> float:
    return lib().orc_bf16_round(v)


def gelu(z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    g, dg = np.empty_like(z), np.empty_like(z)
    lib().orc_gelu(z.size, _p(z), _p(g), _p(dg))
    return g, dg


def mlp_forward(w, X, bf16=True):
    """FP64 forward of the correction network: X[m][128] -> Y[m][32]."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    m = X.shape[0]
    Y = np.empty((m, MLP_KOUT))
    lib().orc_mlp_forward(_p(w), m, _p(X), int(bf16), _p(Y))
    return Y


def mlp_grad(w, X, T, bf16=True):
    """(loss, dL/dw) of L = sum_r T[r,3] |y[r,:3] - T[r,:3]|^2 / m (T is m x 4:
    targets and a per-sample weight)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    g = np.empty(mlp_num_params())
    loss = lib().orc_mlp_grad(_p(w), X.shape[0], _p(X), _p(T), int(bf16), _p(g))
    return loss, g


# ---------------------------------------------------------------- cosmology
def cosmo(H, omega_m=1.0, omega_r=0.0, omega_lam=0.0) -> Cosmo:
    return Cosmo(H, omega_m, omega_r, omega_lam)


def cosmos_update(c: Cosmo, adotdot, adot, a, dt0):
    v = [C.c_double(adotdot), C.c_double(adot), C.c_double(a)]
    lib().orc_cosmos_update(C.byref(c), C.byref(v[0]), C.byref(v[1]), C.byref(v[2]), dt0)
    return v[0].value, v[1].value, v[2].value


def cosmos_dadt(c: Cosmo, a):
    return lib().orc_cosmos_dadt(C.byref(c), a)


def cosmos_conformal_time(c: Cosmo, a0, a1):
    return lib().orc_cosmos_conformal_time(C.byref(c), a0, a1)


def cosmos_time(c: Cosmo, a0, a1):
    return lib().orc_cosmos_time(C.byref(c), a0, a1)


def options_derive(ngrid, mass_res, hubble, omega_m, theta):
    """Derived options (options.cpp:88-95): a dict of GM, omega_r, omega_lam
    (FP32 values), code_to_g, code_to_cm, code_to_s and the H of
    cosmology.cpp:15 (FP64)."""
    out = np.zeros(7, np.float64)
    lib().orc_options_derive(ngrid, mass_res, hubble, omega_m, theta, _p(out))
    keys = ("GM", "code_to_g", "code_to_cm", "code_to_s", "omega_r", "omega_lam", "H")
    return {k: float(v) for k, v in zip(keys, out)}
