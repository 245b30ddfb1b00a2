"""ctypes declarations of libct2hip.so (the C ABI in include/ct2hip.h).

The library is built in-tree by ``ct2_amd/build.py``.  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libct2hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ct2hip.h")

OK = 0
ERR_INVALID = 1
ERR_STATE = 2
ERR_NOMEM = 3
ERR_HIP = 4
ERR_NODEVICE = 5

KERNEL_PLUMMER = 0
KERNEL_CT2 = 1
KICK_NAIVE = 0x1
KICK_STORE = 0x2
KICK_ORDERED = 0x4
KICK_SYMMETRIC = 0x8
PATH_NONE = 0
PATH_DIRECT = 1
PATH_PM = 2
PATH_PM_NEAR = 3
PATH_PM_MLP = 4
PATH_MM = 5


class Ct2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ct2 error {code}: {msg}")
        self.code = code


class ForceOptions(C.Structure):
    _fields_ = [("hsoft", C.c_float), ("GM", C.c_float), ("kernel", C.c_int)]


class KickParams(C.Structure):
    """kick_params, kick.hpp:10-20."""

    _fields_ = [("ascending", C.c_int), ("descending", C.c_int), ("top", C.c_int),
                ("first", C.c_int), ("rung", C.c_int), ("t0", C.c_float), ("eta", C.c_float),
                ("max_dt", C.c_float), ("scale", C.c_float)]


class Cosmology(C.Structure):
    _fields_ = [("H", C.c_double), ("omega_m", C.c_double), ("omega_r", C.c_double), ("omega_lam", C.c_double)]


class Options(C.Structure):
    """ct2_options: the options block of options.hpp:14-32."""

    _fields_ = [("bwidth", C.c_float), ("code_to_cm", C.c_double), ("code_to_g", C.c_double),
                ("code_to_s", C.c_double)] + \
        [(k, C.c_float) for k in ("eta", "GM", "hsoft", "hubble", "omega_lam", "omega_m", "omega_r", "mass_res")] + \
        [("Ngrid", C.c_uint64), ("nparts", C.c_uint64), ("ppcell", C.c_int), ("Theta", C.c_float)]


class ParticlePtrs(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("vx", C.c_void_p),
                ("vy", C.c_void_p), ("vz", C.c_void_p), ("rung", C.c_void_p), ("n", C.c_size_t)]


_P = C.c_void_p
_SIGS = {
    "ct2_last_error": ([], C.c_char_p),
    "ct2_init": ([C.c_int], C.c_int),
    "ct2_finalize": ([], C.c_int),
    "ct2_get_stream": ([C.POINTER(C.c_void_p)], C.c_int),
    "ct2_synchronize": ([], C.c_int),
    "ct2_set_force_options": ([C.POINTER(ForceOptions)], C.c_int),
    "ct2_particles_alloc": ([C.c_size_t], C.c_int),
    "ct2_particle_count": ([], C.c_size_t),
    "ct2_particles_upload": ([_P] * 7, C.c_int),
    "ct2_particles_download": ([_P] * 7, C.c_int),
    "ct2_particles_device_ptrs": ([C.POINTER(ParticlePtrs)], C.c_int),
    "ct2_host_alloc": ([C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "ct2_host_free": ([C.c_void_p], C.c_int),
    "ct2_kick": ([C.POINTER(KickParams), C.c_uint, C.POINTER(C.c_int)], C.c_int),
    "ct2_kick_noalloc": ([C.POINTER(KickParams), C.c_uint, C.POINTER(C.c_int)], C.c_int),
    "ct2_mm_init": ([C.c_int, C.c_int], C.c_int),
    "ct2_mm_cells_for": ([C.c_size_t, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "ct2_mm_forces": ([], C.c_int),
    "ct2_kick_mm": ([C.POINTER(KickParams), C.c_uint, C.POINTER(C.c_int)], C.c_int),
    "ct2_mm_table_download": ([_P], C.c_int),
    "ct2_forces_download": ([_P] * 4, C.c_int),
    "ct2_kick_sample_fp64": ([_P, C.c_size_t, C.c_int, _P, _P, _P, _P], C.c_int),
    "ct2_drift": ([C.c_float, C.c_float], C.c_int),
    "ct2_pm_init": ([C.c_int, C.c_double], C.c_int),
    "ct2_pm_forces": ([], C.c_int),
    "ct2_pm_near": ([C.c_double], C.c_int),
    "ct2_particle_sort_by_rung": ([C.c_int], C.c_int),
    "ct2_particle_push_rungs": ([], C.c_int),
    "ct2_particle_pop_rungs": ([], C.c_int),
    "ct2_particle_current_range": ([_P, _P], C.c_int),
    "ct2_particle_rung_counts": ([_P, C.c_int, _P], C.c_int),
    "ct2_particle_set_minrung": ([C.c_int], C.c_int),
    "ct2_particles_to_grid": ([C.c_int, _P], C.c_int),
    "ct2_debug_fail_allocations": ([C.c_int, C.c_int], C.c_int),
    "ct2_pm_near_cells": ([], C.c_int),
    "ct2_near_kernel_table": ([C.c_double, C.c_double, C.c_int, _P, _P], C.c_int),
    "ct2_pm_assignment": ([C.c_int], C.c_int),
    "ct2_kick_pm": ([C.POINTER(KickParams), C.c_uint, C.POINTER(C.c_int)], C.c_int),
    "ct2_pm_density_download": ([_P], C.c_int),
    "ct2_pm_grid_download": ([C.c_int, _P], C.c_int),
    "ct2_mlp_num_params": ([], C.c_size_t),
    "ct2_mlp_init": ([C.c_uint64], C.c_int),
    "ct2_mlp_get_weights": ([_P, C.c_size_t], C.c_int),
    "ct2_mlp_set_weights": ([_P, C.c_size_t], C.c_int),
    "ct2_mlp_forward": ([_P, C.c_size_t, _P], C.c_int),
    "ct2_mlp_grad": ([_P, _P, C.c_size_t, _P, _P], C.c_int),
    "ct2_mlp_features": ([_P, C.c_size_t, _P], C.c_int),
    "ct2_mlp_data_reset": ([C.c_size_t], C.c_int),
    "ct2_mlp_data_size": ([], C.c_size_t),
    "ct2_mlp_data_append": ([_P, C.c_size_t], C.c_int),
    "ct2_mlp_train": ([C.c_int, C.c_int, C.c_float, C.c_uint64, _P], C.c_int),
    "ct2_pm_mlp_forces": ([], C.c_int),
    "ct2_mlp_set_mesh": ([C.c_double, C.c_int, C.c_double, C.c_double], C.c_int),
    "ct2_last_force_path": ([], C.c_int),
    "ct2_kick_pm_mlp": ([C.POINTER(KickParams), C.c_uint, C.POINTER(C.c_int)], C.c_int),
    "ct2_fft3d_c2c": ([_P, C.c_int, C.c_int], C.c_int),
    "ct2_fft3d_c2c_f64": ([_P, C.c_int, C.c_int], C.c_int),
    "ct2_fft_complex_1d": ([_P, C.c_int, C.c_int, C.c_int, C.c_size_t], C.c_int),
    "ct2_fft_real_1d": ([_P, C.c_int, C.c_int, C.c_int, C.c_size_t], C.c_int),
    "ct2_scramble": ([_P, C.c_int, C.c_int, C.c_int], C.c_int),
    "ct2_transpose_yz": ([_P, C.c_int], C.c_int),
    "ct2_transpose_r2c": ([_P, _P, C.c_int], C.c_int),
    "ct2_transpose_c2r": ([_P, _P, C.c_int], C.c_int),
    "ct2_fft_r2c_3d": ([_P, _P, C.c_int], C.c_int),
    "ct2_fft_c2r_3d": ([_P, _P, C.c_int], C.c_int),
    "ct2_twiddles": ([C.c_int, _P, _P, _P], C.c_int),
    "ct2_dev_alloc": ([C.c_size_t, C.POINTER(C.c_void_p)], C.c_int),
    "ct2_dev_free": ([_P], C.c_int),
    "ct2_memcpy_h2d": ([_P, _P, C.c_size_t], C.c_int),
    "ct2_memcpy_d2h": ([_P, _P, C.c_size_t], C.c_int),
    "ct2_ic_lattice": ([C.c_int, C.c_double, C.c_uint64, _P, _P, _P], C.c_int),
    "ct2_ic_grf": ([C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_double], C.c_int),
    "ct2_cosmos_update": ([C.POINTER(Cosmology), _P, _P, _P, C.c_double], C.c_int),
    "ct2_cosmos_dadt": ([C.POINTER(Cosmology), C.c_double], C.c_double),
    "ct2_cosmos_conformal_time": ([C.POINTER(Cosmology), C.c_double, C.c_double], C.c_double),
    "ct2_cosmos_time": ([C.POINTER(Cosmology), C.c_double, C.c_double], C.c_double),
    "ct2_cosmos_kick_factor": ([C.POINTER(Cosmology), C.c_double, C.c_double], C.c_double),
    "ct2_cosmos_drift_factor": ([C.POINTER(Cosmology), C.c_double, C.c_double], C.c_double),
    "ct2_options_defaults": ([C.POINTER(Options)], C.c_int),
    "ct2_options_derive": ([C.POINTER(Options)], C.c_int),
    "ct2_options_cosmology": ([C.POINTER(Options), C.POINTER(Cosmology)], C.c_int),
    "ct2_ic_plummer_mix": ([C.c_size_t, C.c_size_t, C.c_int, C.c_double, C.c_double, C.c_uint64,
                            _P, _P, _P], C.c_int),
}

_lib = None


def load():
    """Loads libct2hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`"
                               " (there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise Ct2Error(rc, load().ct2_last_error().decode(errors="replace"))


def declared_symbols() -> list[str]:
    """Function names declared in include/ct2hip.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ct2_[a-z0-9_]+)\s*\(", text)))
