"""Host-side parts of the learned warm start (no GPU): the product's network
definition agrees with the oracle's, and the weight file round-trips."""

from __future__ import annotations

import numpy as np
import pytest

import net_oracle as no
from darcy import learned


@pytest.mark.parametrize("sub", [1, 2])
def test_layout_and_init_match_oracle(sub):
    assert learned.param_count(sub) == no.layout(sub)[1]
    np.testing.assert_array_equal(learned.init_params(5, sub), no.init_params(5, sub))
    assert learned.sub_of_count(learned.param_count(sub)) == sub


@pytest.mark.parametrize("sub", [1, 2])
def test_weight_file_roundtrip(tmp_path, sub):
    p = learned.init_params(2, sub)
    p[-300:] = np.random.default_rng(0).normal(size=300).astype(np.float32)  # head + last biases in FP32
    f = tmp_path / "w.npz"
    learned.save_params(str(f), p, epochs=1)
    q, got_sub = learned.load_weights(str(f))
    assert got_sub == sub
    idx = learned._fp32_index(sub)
    np.testing.assert_array_equal(q[idx], p[idx])  # biases and head exact
    conv = np.setdiff1d(np.arange(p.size), idx)
    np.testing.assert_array_equal(q[conv], no.bf16(p[conv]).astype(np.float32))  # conv weights as bf16
    with np.load(f, allow_pickle=False) as z:
        assert set(z.files) >= {"weights_bf16", "fp32_index", "fp32_values"}


def test_committed_weights_load():
    p, sub = learned.load_weights()
    assert p.size == learned.param_count(sub) and np.all(np.isfinite(p))
