"""Pins the CPU oracle of the warm-start network (oracle/net_oracle.py).

The oracle is what the GPU network is checked against, so it is pinned here
by things that do not restate it: torch's bf16 conversion, a direct
seven-loop convolution, the adjoint identities of the convolution, central
finite differences of its own FP64 loss, and torch autograd in FP64 on the
same network.  A planted mistake in the oracle's backward pass must fail the
finite-difference check.
"""

from __future__ import annotations

import numpy as np
import pytest

import net_oracle as no
from conftest import rel_l2, two_phase


def _batch(n: int, seeds: list[int]) -> np.ndarray:
    return np.stack([two_phase(n, seed=s) for s in seeds])


def _target(a: np.ndarray) -> np.ndarray:
    # A smooth positive field of the solution's size (the loss only needs a target).
    n = a.shape[1]
    pos = no.position(n)
    return 0.012 * pos[None] * (6.0 / a) ** 0.3


def test_layout_counts():
    lay, total = no.layout()
    assert len(lay) == 2 * no.LEVELS
    assert total == sum(int(np.prod(s)) + s[0] for _, _, s in lay)
    assert total == 4317921
    assert lay[0][2] == (32, 3, 3, 8) and lay[-1][2] == (1, 3, 3, 32)
    # D_1 is 32 -> 64 at stride 2; U_0 maps 64 channels to 32
    assert no.conv_specs()[1] == (32, 64, 2, 0) and no.conv_specs()[-1] == (64, 32, 1, 1)
    # The half-resolution network differs only in its head: 32 -> 5.
    lay2, total2 = no.layout(2)
    assert lay2[-1][2] == (5, 3, 3, 32) and total2 == total + 4 * 9 * 32 + 4
    assert no.sub_of(np.zeros(total)) == 1 and no.sub_of(np.zeros(total2)) == 2


def test_pixel_shuffle_is_an_orthogonal_reindexing():
    """Shuffle and unshuffle are inverse permutations (so each is the other's adjoint),
    and channel di*2+dj of pixel (I, J) lands on cell (2I+di, 2J+dj)."""
    rng = np.random.default_rng(4)
    y = rng.normal(size=(2, 3, 3, 4))
    f = no.pixel_shuffle(y, 2)
    np.testing.assert_array_equal(no.pixel_unshuffle(f, 2), y)
    assert f[1, 2 * 2 + 1, 2 * 0 + 0] == y[1, 2, 0, 2]  # I=2, J=0, di=1, dj=0
    assert f[0, 2 * 1 + 0, 2 * 2 + 1] == y[0, 1, 2, 1]  # I=1, J=2, di=0, dj=1
    d = rng.normal(size=f.shape)
    assert abs(np.sum(f * d) - np.sum(y * no.pixel_unshuffle(d, 2))) < 1e-12


def test_sub2_output_map_against_torch_and_its_adjoint():
    """Channel 0 interpolated as torch's bilinear x2 upsampling (align_corners
    False, edges clamped), the detail channels as torch's pixel shuffle; the
    adjoint identity <out(yc), d> = <yc, adjoint(d)>; a constant channel 0
    interpolates to the constant."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(6)
    yc = rng.normal(size=(2, 5, 5, 5))
    t = torch.tensor(yc).permute(0, 3, 1, 2)
    want = (F.interpolate(t[:, :1], scale_factor=2, mode="bilinear", align_corners=False)
            + F.pixel_shuffle(t[:, 1:], 2))[:, 0].numpy()
    np.testing.assert_allclose(no.sub2_out(yc), want, rtol=1e-13, atol=1e-13)
    d = rng.normal(size=(2, 10, 10))
    assert abs(np.sum(no.sub2_out(yc) * d) - np.sum(yc * no.sub2_adjoint(d))) < 1e-11
    c = np.zeros((1, 4, 4, 5))
    c[..., 0] = 2.5
    np.testing.assert_allclose(no.sub2_out(c), 2.5, rtol=1e-15)


def test_half_resolution_features_are_block_values():
    """Channel r of pixel (I, J) is the normalised ln a of cell (2I + r // 2, 2J + r % 2);
    the block mean of a constant field's position is its mean; a sub=1 call is
    the full-resolution map."""
    a = _batch(64, [3])
    x2 = no.features(a, 2)
    x1 = no.features(a, 1)
    assert x2.shape == (1, 32, 32, 8) and x1.shape == (1, 64, 64, 8)
    for r in range(4):
        np.testing.assert_array_equal(x2[..., r], x1[:, r // 2::2, r % 2::2, 0])
    assert np.all(x2[..., 4] == 1.0) and np.all(x2[..., 6:] == 0.0)
    blocks = x1[..., 2].reshape(1, 32, 2, 32, 2).mean(axis=(2, 4))
    np.testing.assert_allclose(x2[..., 5], blocks, rtol=1e-6)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.normal(size=20000) * 10.0 ** rng.integers(-8, 8, size=20000),
        # exact ties between two bf16 values, both parities
        (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32).astype(np.float64),
        np.array([0.0, -0.0, 1.0, -1.0, 3.0e38, 1e-40]),
    ]).astype(np.float32)
    ours = no.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(ours, ref)
    np.testing.assert_array_equal(no.from_bits(no.bf16_bits(x)), ref)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_direct_sum(stride):
    rng = np.random.default_rng(2 + stride)
    x = rng.normal(size=(2, 6, 4, 3))
    w = rng.normal(size=(5, 3, 3, 3))
    b = rng.normal(size=5)
    np.testing.assert_allclose(no.conv_forward(x, w, b, stride), no.conv_direct(x, w, b, stride), rtol=1e-12,
                               atol=1e-12)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_adjoint_identities(stride):
    rng = np.random.default_rng(7 + stride)
    x = rng.normal(size=(2, 8, 8, 4))
    w = rng.normal(size=(6, 3, 3, 4))
    z = no.conv_forward(x, w, np.zeros(6), stride)
    dz = rng.normal(size=z.shape)
    lhs = float(np.sum(z * dz))
    dx = no.conv_backward_data(dz, w, x.shape, stride)
    dw, db = no.conv_backward_weight(x, dz, stride)
    assert abs(lhs - float(np.sum(x * dx))) <= 1e-10 * abs(lhs) + 1e-10
    assert abs(lhs - float(np.sum(w * dw))) <= 1e-10 * abs(lhs) + 1e-10
    np.testing.assert_allclose(db, dz.sum(axis=(0, 1, 2)), rtol=1e-12)


def _fd_errors(net_cls, params, a, u, picks):
    net = net_cls(params)
    _, g = net.gradient(a, u)
    # Relative error, with the denominator floored at 1e-5 of the largest
    # gradient: a pick whose gradient is at round-off level compares noise.
    floor = 1e-5 * float(np.abs(g).max())
    errs = []
    for i in picks:
        h = 1e-5 * max(1.0, abs(float(params[i])))
        pp, pm = params.astype(np.float64).copy(), params.astype(np.float64).copy()
        pp[i] += h
        pm[i] -= h
        fd = (net_cls(pp).loss(a, u)[0] - net_cls(pm).loss(a, u)[0]) / (2 * h)
        errs.append(abs(fd - g[i]) / max(abs(fd), abs(g[i]), floor, 1e-12))
    return np.array(errs), g


def _picks(params, rng, per_layer=3):
    lay, _ = no.layout(no.sub_of(params))
    out = []
    for w, b, shape in lay:
        out.extend(rng.integers(w, b, size=per_layer).tolist())  # weights
        out.append(b + int(rng.integers(0, shape[0])))  # a bias
    return out


def _fd_problem(sub=1):
    rng = np.random.default_rng(11)
    params = no.init_params(3, sub)
    params[no.layout()[0][0][1]:no.layout()[0][0][1] + 32] = rng.normal(size=32) * 0.1  # nonzero biases
    a = _batch(64 * sub, [5])
    return rng, params, a, _target(a)


@pytest.mark.parametrize("sub", [1, 2])
def test_gradient_matches_finite_differences(sub):
    rng, params, a, u = _fd_problem(sub)
    # Only channels 0..sub^2+1 of the first conv see nonzero input; pick weights there.
    picks = [p for p in _picks(params, rng) if not (p < no.layout()[0][0][1] and p % 8 >= sub * sub + 2)]
    if sub == 2:  # every head channel
        w_head = no.layout(2)[0][-1][0]
        picks += [w_head + r * 288 + 17 for r in range(5)]
    errs, g = _fd_errors(no.Net, params, a, u, picks)
    assert np.max(errs) < 1e-5, errs
    assert np.count_nonzero(g) > 0.5 * g.size


def test_planted_mistake_fails_finite_differences():
    """A wrong adjoint of the upsampling (mean instead of sum) in the oracle must be caught."""

    class Broken(no.Net):
        def gradient(self, a, u):
            saved = no.down_sum
            try:
                # Wrong adjoint of the nearest upsampling: mean instead of sum.
                no.down_sum = lambda d: saved(d) / 4.0
                return super().gradient(a, u)
            finally:
                no.down_sum = saved

    rng, params, a, u = _fd_problem()
    lay, _ = no.layout()
    u0 = 2 * no.LEVELS - 2  # U_0: its weights sit behind the upsampling
    picks = list(range(lay[u0][0], lay[u0][0] + 5))
    errs, _ = _fd_errors(Broken, params, a, u, picks)
    assert np.max(errs) > 0.1


@pytest.mark.parametrize("sub", [1, 2])
def test_gradient_matches_torch_autograd(sub):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng, params, a, u = _fd_problem(sub)
    n = 64 * sub
    a = _batch(n, [5, 6])
    u = _target(a)
    loss_np, g_np = no.Net(params).gradient(a, u)

    lay, _ = no.layout(sub)
    ts = []
    for w, b, shape in lay:
        ts.append(torch.tensor(params[w:b].reshape(shape), dtype=torch.float64, requires_grad=True))
        ts.append(torch.tensor(params[b:b + shape[0]], dtype=torch.float64, requires_grad=True))

    def conv(x, i, stride):  # x: NCHW; weight [co][ky][kx][ci] -> [co][ci][ky][kx]
        return F.conv2d(x, ts[2 * i].permute(0, 3, 1, 2), ts[2 * i + 1], stride=stride, padding=1)

    x = torch.tensor(no.features(a, sub), dtype=torch.float64).permute(0, 3, 1, 2)
    acts = []
    for l in range(no.LEVELS):
        x = F.gelu(conv(x, l, 1 if l == 0 else 2), approximate="tanh")
        acts.append(x)
    for i in range(no.LEVELS - 2, -1, -1):
        c = conv(x, 2 * no.LEVELS - 2 - i, 1)
        x = F.gelu(F.interpolate(c, scale_factor=2, mode="nearest") + acts[i], approximate="tanh")
    yc = conv(x, 2 * no.LEVELS - 1, 1)
    if sub == 1:
        y = yc[:, 0]
    else:  # torch's bilinear upsampling and pixel shuffle
        y = (F.interpolate(yc[:, :1], scale_factor=2, mode="bilinear", align_corners=False)
             + F.pixel_shuffle(yc[:, 1:], 2))[:, 0]
    pos = torch.tensor(no.position(n), dtype=torch.float64)
    t = torch.tensor(u / no.SCALE, dtype=torch.float64)
    e = pos * y - t
    at = torch.tensor(a, dtype=torch.float64)
    fe = 2 * at[:, :, 1:] * at[:, :, :-1] / (at[:, :, 1:] + at[:, :, :-1])
    fn = 2 * at[:, 1:, :] * at[:, :-1, :] / (at[:, 1:, :] + at[:, :-1, :])
    energy = (torch.sum(fe * (e[:, :, 1:] - e[:, :, :-1]) ** 2) + torch.sum(fn * (e[:, 1:, :] - e[:, :-1, :]) ** 2)
              + torch.sum(at[:, :, 0] * e[:, :, 0] ** 2) + torch.sum(at[:, :, -1] * e[:, :, -1] ** 2)
              + torch.sum(at[:, 0, :] * e[:, 0, :] ** 2) + torch.sum(at[:, -1, :] * e[:, -1, :] ** 2))
    loss = (torch.sum(e ** 2) + no.ENERGY_WEIGHT * energy) / e.numel()
    loss.backward()
    g_t = torch.cat([p.grad.reshape(-1) for p in ts]).numpy()
    assert abs(loss.item() - loss_np) <= 1e-12 * abs(loss_np)
    assert rel_l2(g_np, g_t) < 1e-10


def test_loss_operator_is_the_solver_operator(oracle):
    """The loss's energy term uses the finite-volume operator of the solve."""
    rng = np.random.default_rng(21)
    a = _batch(64, [7, 8])
    x = rng.normal(size=a.shape)
    for ai, xi, yi in zip(a, x, no.operator(a, x)):
        np.testing.assert_allclose(yi, oracle.apply_operator(ai, xi), rtol=1e-12, atol=1e-12)


def test_emulated_rounding_stays_close_to_exact():
    rng, params, a, u = _fd_problem()
    l0, g0 = no.Net(params).gradient(a, u)
    l1, g1 = no.Net(params, emulate=True).gradient(a, u)
    assert abs(l1 - l0) < 2e-2 * l0
    lay, _ = no.layout()
    for w, b, shape in lay:
        assert rel_l2(g1[w:b + shape[0]], g0[w:b + shape[0]]) < 5e-2
