"""The cubic pieces the near-field kernel evaluates for its filtered part
(ct2_near_kernel_table): against the closed forms in FP64 (scipy's erf),
their r -> 0 limits, and continuity between pieces."""
import numpy as np
import pytest
from scipy.special import erf

from ct2_amd import api

C = 2 / np.sqrt(np.pi)


def closed(r2, rs):
    u = np.sqrt(r2) / (2 * rs)
    kl = (erf(u) - C * u * np.exp(-u * u)) / u ** 3 / (2 * rs) ** 3
    pl = erf(u) / u / (2 * rs)
    return kl, pl


def evaluate(k4, p4, r2, rc):
    """The kernel's evaluation order: t = r^2 nt / rc^2 in FP32, piece min(int(t), nt - 1)."""
    nt = k4.shape[0]
    t = (r2.astype(np.float32) * np.float32(nt / rc ** 2)).astype(np.float32)
    i = np.minimum(t.astype(np.int64), nt - 1)
    f = (t - i.astype(np.float32)).astype(np.float64)
    ev = lambda c: ((c[i, 3] * f + c[i, 2]) * f + c[i, 1]) * f + c[i, 0]
    return ev(k4.astype(np.float64)), ev(p4.astype(np.float64))


@pytest.mark.parametrize("rs_cells,ng,rcut_rs", [(1.5, 256, 5.5), (2.5, 512, 5.5), (4.0, 512, 5.5), (1.0, 64, 6.0)])
def test_pieces_match_closed_form(rs_cells, ng, rcut_rs):
    rs = rs_cells / ng * 2.0 ** 32
    rc = rcut_rs * rs
    k4, p4 = api.near_kernel_table(rs, rc)
    rng = np.random.default_rng(3)
    r2 = np.concatenate([rng.uniform(0.0, 1.0, 20000) ** 2, rng.uniform(0.0, 1.0, 20000)]) * rc ** 2
    r2 = r2[r2 > (1e-3 * rs) ** 2]
    kl, pl = evaluate(k4, p4, r2, rc)
    wk, wp = closed(r2, rs)
    # within a few FP32 roundings (of the coefficients, t and f) of the
    # filtered kernel itself; the cubic fit alone is within 2e-8
    assert np.max(np.abs(kl - wk) / wk) < 4e-7
    assert np.max(np.abs(pl - wp) / wp) < 4e-7


def test_limits_at_zero_separation():
    rs = 1.5 / 256 * 2.0 ** 32
    k4, p4 = api.near_kernel_table(rs, 5.5 * rs)
    assert k4[0, 0] == pytest.approx(C * (2 / 3) / (2 * rs) ** 3, rel=1e-7)
    assert p4[0, 0] == pytest.approx(C / (2 * rs), rel=1e-7)


def test_pieces_are_continuous():
    rs = 2.0 / 128 * 2.0 ** 32
    k4, p4 = api.near_kernel_table(rs, 5.5 * rs)
    for c in (k4.astype(np.float64), p4.astype(np.float64)):
        end = c.sum(axis=1)[:-1]  # piece i at f = 1
        start = c[1:, 0]          # piece i + 1 at f = 0
        assert np.max(np.abs(end - start) / np.abs(start)) < 2.5e-7  # FP32 coefficients


def test_bad_arguments():
    from ct2_amd._lib import Ct2Error
    with pytest.raises(Ct2Error):
        api.near_kernel_table(0.0, 1.0)
    with pytest.raises(Ct2Error):
        api.near_kernel_table(1.0, 1.0, 0)
