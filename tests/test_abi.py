"""The C ABI surface: the library loads, exports every symbol the header
declares, and rejects calls made in the wrong state or with bad arguments
(no GPU needed; nothing here launches a kernel)."""
import ctypes as C

import numpy as np
import pytest

from ct2_amd import _lib as L


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = L.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L._SIGS), "ctypes declarations out of sync with the header"


def test_calls_before_init_report_state_errors():
    lib = L.load()
    assert lib.ct2_particles_alloc(10) == L.ERR_STATE
    assert b"ct2_init" in lib.ct2_last_error()
    assert lib.ct2_drift(1.0, 1.0) == L.ERR_STATE
    p = L.KickParams()
    assert lib.ct2_kick(C.byref(p), 0, None) == L.ERR_STATE
    assert lib.ct2_fft3d_c2c(None, 64, 1) == L.ERR_STATE


def test_option_validation_is_host_side():
    lib = L.load()
    bad = [L.ForceOptions(0.0, 1.0, 0), L.ForceOptions(0.01, -1.0, 0), L.ForceOptions(0.01, 1.0, 7)]
    for o in bad:
        assert lib.ct2_set_force_options(C.byref(o)) == L.ERR_INVALID
    assert lib.ct2_set_force_options(C.byref(L.ForceOptions(0.02, 1.0, 1))) == L.OK


def test_mm_cells_for_is_host_side():
    """The multipole mesh's cell grid: the largest power of two in [8, 256]
    with at least the given particles per cell on average."""
    from ct2_amd import api
    assert api.mm_cells_for(2 ** 21) == 64 and api.mm_cells_for(2 ** 24) == 128
    assert api.mm_cells_for(2 ** 21, 64) == 32 and api.mm_cells_for(2 ** 21, 1) == 128
    assert api.mm_cells_for(100) == 8 and api.mm_cells_for(0) == 8
    assert api.mm_cells_for(2 ** 31) == 256
    assert api.mm_cells_for(8 * 64 ** 3 - 1) == 32  # one short of 8 per cell on 64^3
    nc = C.c_int(0)
    assert L.load().ct2_mm_cells_for(1000, 0, C.byref(nc)) == L.ERR_INVALID
    assert L.load().ct2_mm_cells_for(1000, 8, None) == L.ERR_INVALID


def test_host_alloc_argument_checks_are_host_side():
    """ct2_host_alloc: a null output is rejected and 0 bytes give NULL
    without touching the runtime; ct2_host_free takes NULL."""
    lib = L.load()
    assert lib.ct2_host_alloc(16, None) == L.ERR_INVALID
    p = C.c_void_p(1234)
    assert lib.ct2_host_alloc(0, C.byref(p)) == L.OK and not p.value
    assert lib.ct2_host_free(None) == L.OK
