"""The erfc polynomial of the near-field kernel (ct2_amd/csrc/near_erfc.h),
checked against scipy's erfc: the FP32 Horner evaluation the kernel performs
must stay within a few units in the last place of 1 on the fitted interval,
and the clamp beyond it must stay below FP32 resolution."""
import os
import re

import numpy as np
from scipy.special import erfc

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ct2_amd", "csrc", "near_erfc.h")


def coefficients():
    text = open(HDR).read()
    body = re.search(r"NEAR_ERFC_C\[[^]]*\]\s*=\s*\{([^}]*)\}", text).group(1)
    return np.array([float(v.strip().rstrip("f")) for v in body.split(",")], np.float32)


def kernel_erfc(u):
    """The kernel's arithmetic in FP32 (fmaf emulated by FP64 then rounding)."""
    C = coefficients()
    u = np.asarray(u, np.float32)
    s = np.minimum(np.float32(u.astype(np.float64) * 0.5 - 1.25), np.float32(1.0))
    g = np.full_like(u, C[-1])
    for c in C[-2::-1]:
        g = (g.astype(np.float64) * s + c).astype(np.float32)
    ex = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    return ex.astype(np.float64) * g


def test_fit_on_interval():
    u = np.linspace(0.5, 4.5, 200001)
    err = np.abs(kernel_erfc(u) - erfc(np.asarray(u, np.float32).astype(np.float64)))
    assert err.max() < 1.5e-7, err.max()


def test_clamp_beyond_interval():
    u = np.linspace(4.5, 12.0, 10001)
    err = np.abs(kernel_erfc(u) - erfc(np.asarray(u, np.float32).astype(np.float64)))
    assert err.max() < 1e-9


def test_header_matches_generator():
    import subprocess
    import sys
    gen = os.path.join(os.path.dirname(HDR), "fit_erfc.py")
    out = subprocess.run([sys.executable, gen], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert out == open(HDR).read()
