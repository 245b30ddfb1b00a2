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
