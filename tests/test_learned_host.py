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


def test_training_rolls_back_a_loss_spike(monkeypatch):
    """A loss spike above 10x the best epoch's loss restores that epoch's
    parameters (which also restarts Adam) and halves the rate; the host
    logic of learned.train, driven by a stand-in network."""
    rates, restored = [], []
    epoch_losses = iter([1.0, 0.5, 0.4, 9.0, 0.3, 0.2])  # epoch 4 spikes

    class FakeNet:
        def __init__(self, ctx, n, batch, train=False, sub=1):
            self.params = np.zeros(3, dtype=np.float32)
            self.loss = None

        def set_params(self, p):
            restored.append(np.array(p))
            self.params = np.array(p, dtype=np.float32)

        def get_params(self):
            return self.params.copy()

        def train_step(self, a, u, rate, want_loss=True):
            if self.loss is None:
                self.loss = next(epoch_losses)
            rates.append(rate)
            self.params = self.params + 1.0  # every step moves the weights
            return self.loss

        def close(self):
            pass

    class Data:
        n, count, a, u = 64, 4, [0, 1, 2, 3], [0, 1, 2, 3]

    monkeypatch.setattr(learned.capi, "Net", FakeNet)
    monkeypatch.setattr(learned, "init_params", lambda seed, sub=1: np.zeros(3, dtype=np.float32))
    net_holder = {}
    orig_train_step = FakeNet.train_step

    def train_step(self, a, u, rate, want_loss=True):
        net_holder["net"] = self
        out = orig_train_step(self, a, u, rate, want_loss)
        return out

    monkeypatch.setattr(FakeNet, "train_step", train_step)

    def new_epoch_hook(entry_json):
        net_holder["net"].loss = None  # the next epoch draws its loss

    params, history = learned.train(None, Data(), epochs=6, batch=2, lr=1.0, log=new_epoch_hook)
    assert [h.get("rolled_back", False) for h in history] == [False, False, False, True, False, False]
    assert history[3]["rate_scale"] == 0.5
    # the rollback restored the parameters kept after epoch 3 (the best so far): 3 epochs x 2 steps
    assert np.array_equal(restored[-1], np.full(3, 6.0, dtype=np.float32))
    # rates after the rollback are half of what the schedule gives
    # (lr 1, one warm-up step, cosine over 12 steps)
    sched = [0.5 * (1.0 + np.cos(np.pi * k / 12)) for k in range(12)]
    np.testing.assert_allclose(rates[:8], sched[:8], rtol=1e-12)
    np.testing.assert_allclose(rates[8:], 0.5 * np.array(sched[8:]), rtol=1e-12)
