"""Host-side parts of the learned warm start (no GPU): the product's network
definition agrees with the oracle's, and the weight file round-trips."""

from __future__ import annotations

import numpy as np

import net_oracle as no
from darcy import learned


def test_layout_and_init_match_oracle():
    assert learned.param_count() == no.layout()[1]
    np.testing.assert_array_equal(learned.init_params(5), no.init_params(5))


def test_weight_file_roundtrip(tmp_path):
    p = learned.init_params(2)
    p[-300:] = np.random.default_rng(0).normal(size=300).astype(np.float32)  # head + last biases in FP32
    f = tmp_path / "w.npz"
    learned.save_params(str(f), p, epochs=1)
    q = learned.load_params(str(f))
    idx = learned._fp32_index()
    np.testing.assert_array_equal(q[idx], p[idx])  # biases and head exact
    conv = np.setdiff1d(np.arange(p.size), idx)
    np.testing.assert_array_equal(q[conv], no.bf16(p[conv]).astype(np.float32))  # conv weights as bf16
    with np.load(f, allow_pickle=False) as z:
        assert set(z.files) >= {"weights_bf16", "fp32_index", "fp32_values"}


def test_committed_weights_load():
    p = learned.load_params()
    assert p.size == learned.param_count() and np.all(np.isfinite(p))
