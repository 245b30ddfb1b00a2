"""Pins the FP64 oracle of the learned correction: GELU against its closed
form and the erf form, the backward pass against central finite
differences, and bf16 rounding against exact ties."""
import math

import numpy as np
import pytest


def test_gelu_closed_form_and_derivative(orc):
    z = np.linspace(-6, 6, 241)
    g, dg = orc.gelu(z)
    want = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.max(np.abs(g - want)) < 1e-3  # tanh form vs exact erf form
    assert g[120] == 0.0 and abs(g[-1] - 6.0) < 1e-8
    h = 1e-6
    gp, _ = orc.gelu(z + h)
    gm, _ = orc.gelu(z - h)
    assert np.allclose(dg, (gp - gm) / (2 * h), atol=1e-8)



@pytest.mark.parametrize("v,want", [(1.0, 1.0), (1.0 + 2 ** -8, 1.0), (1.0 + 3 * 2 ** -8, 1.0 + 2 ** -6),
                                    (1.0 + 2 ** -7, 1.0 + 2 ** -7), (-(1.0 + 2 ** -8), -1.0), (3.0e38, 3.00405527047391e38)])
def test_bf16_round_cases(orc, v, want):
    r = orc.bf16_round(v)
    assert r == pytest.approx(want, rel=1e-4)


def test_backward_matches_finite_differences(orc):
    rng = np.random.default_rng(0)
    P = orc.mlp_num_params()
    w = rng.standard_normal(P) * 0.05
    m = 5
    X = np.zeros((m, 128))
    X[:, :111] = rng.standard_normal((m, 111))
    T = np.concatenate([rng.standard_normal((m, 3)), rng.uniform(0.5, 2.0, (m, 1))], axis=1)
    loss, g = orc.mlp_grad(w, X, T, bf16=False)
    # offsets of W1, b1, W2, b2, W3, b3 in the flat layout
    offs = [0, 128 * 256, 128 * 256 + 256, 128 * 256 + 256 + 256 * 256, 128 * 256 + 512 + 256 * 256,
            128 * 256 + 512 + 256 * 256 + 256 * 32]
    picks = [offs[0] + 3 * 256 + 7, offs[1] + 11, offs[2] + 100 * 256 + 5, offs[3] + 200, offs[4] + 17 * 32 + 1,
             offs[5] + 2, offs[4] + 9 * 32 + 5]
    h = 1e-6
    for p in picks:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (orc.mlp_grad(wp, X, T, bf16=False)[0] - orc.mlp_grad(wm, X, T, bf16=False)[0]) / (2 * h)
        assert g[p] == pytest.approx(fd, rel=1e-5, abs=1e-10), p
    # outputs 3..31 carry no loss: their weights get no gradient
    assert g[offs[4] + 9 * 32 + 5] == 0.0
    # the loss is the mean squared error of the forward pass
    Y = orc.mlp_forward(w, X, bf16=False)
    want = np.sum(T[:, 3:] * (Y[:, :3] - T[:, :3]) ** 2) / m
    assert loss == pytest.approx(want, rel=1e-12)
