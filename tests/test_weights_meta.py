"""The committed learned-correction weights: each file loads without
executing anything, holds the library's parameter count of finite FP32
values, and records the mesh it was trained to correct (assignment, split,
grid), which bench.py and ct2_amd/stepping.py read back (no GPU needed)."""
import os

import numpy as np
import pytest

from ct2_amd import _lib as L
from ct2_amd.train_mlp import load_weights

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FILES = ["mlp_n128_ng256.npz", "mlp_grf_ng2x.npz"]


@pytest.mark.parametrize("name", FILES)
def test_committed_weights_and_their_mesh(name):
    out = load_weights(os.path.join(ROOT, "ct2_amd", "weights", name))
    assert out is not None, name
    w, meta = out
    assert w.dtype == np.float32 and w.shape == (int(L.load().ct2_mlp_num_params()),)
    assert np.isfinite(w).all() and np.abs(w).max() > 0
    assert meta["order"] in (1, 2) and meta["rs_cells"] > 0 and meta["ng"] == 2 * meta["side"]
    assert meta["near_rcut_rs"] == 0.0  # a correction of the mesh alone
    assert np.isfinite(meta["final_loss"]) and meta["steps"] > 0 and meta["batch"] > 0


def test_missing_weights_are_reported_as_absent():
    assert load_weights(os.path.join(ROOT, "ct2_amd", "weights", "no_such_file.npz")) is None
