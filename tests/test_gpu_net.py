"""The warm-start network's GPU kernels against the CPU network oracle.

Each conv kernel (MFMA implicit GEMM forward, backward data, backward weight)
is compared with the oracle's convolution on bf16 inputs; the whole network's
prediction, loss and parameter gradient with the oracle in bf16-emulating
mode (the GPU then differs only by FP32 accumulation order) and, more
loosely, with the exact FP64 gradient.
"""

from __future__ import annotations

import numpy as np
import pytest

import net_oracle as no
from conftest import rel_l2, two_phase

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("darcy.capi")


@pytest.fixture(scope="module")
def ctx():
    with capi.Context() as c:
        yield c


def _bf(rng, shape, scale=1.0):
    return no.bf16_bits(rng.normal(size=shape) * scale)


def _close_bf16(got_bits, want, label):
    """Elementwise within one bf16 rounding of the FP64 value, up to FP32 accumulation noise."""
    got = no.from_bits(got_bits)
    ref = no.bf16(want)
    tol = 2.0 ** -8 * np.abs(want) + 1e-5 * np.max(np.abs(want))
    bad = np.abs(got - want) > tol
    assert not bad.any(), f"{label}: {bad.sum()} of {bad.size} outside tolerance"
    assert np.mean(got == ref) > 0.99, f"{label}: too few exact bf16 matches"


def test_mfma_operand_layout(ctx):
    rng = np.random.default_rng(0)
    ai = rng.integers(-8, 9, size=(16, 32)).astype(np.float64)
    bi = rng.integers(-8, 9, size=(32, 16)).astype(np.float64)
    c = ctx.mfma_check(no.bf16_bits(ai), no.bf16_bits(bi))
    np.testing.assert_array_equal(c, ai @ bi)  # small integers: exact
    a = no.bf16(rng.normal(size=(16, 32)))
    b = no.bf16(rng.normal(size=(32, 16)))
    np.testing.assert_allclose(ctx.mfma_check(no.bf16_bits(a), no.bf16_bits(b)), a @ b, rtol=1e-5, atol=1e-5)


SHAPES = [  # b, h, w, cin, cout, stride
    (2, 16, 16, 32, 64, 1),
    (2, 16, 16, 8, 32, 1),
    (1, 8, 8, 128, 256, 2),
    (2, 12, 20, 24, 40, 2),
    (4, 10, 6, 16, 24, 1),
    # halo-tiled kernel (stride 1, width a multiple of 32 or 64, Cout 32 / 64)
    (2, 8, 64, 64, 32, 1),
    (1, 16, 32, 32, 64, 1),
    (2, 4, 128, 128, 64, 1),
    (2, 8, 64, 8, 32, 1),  # the 8-channel halo kernel
    (1, 16, 32, 8, 64, 1),
    # stride 2 at the encoder's channel counts: several K chunks and N tiles
    (1, 16, 32, 32, 64, 2),
    (2, 32, 64, 64, 128, 2),
    (1, 16, 32, 128, 256, 2),
]


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_forward(ctx, shape):
    b, h, w, cin, cout, s = shape
    rng = np.random.default_rng(sum(shape))
    x, wt = _bf(rng, (b, h, w, cin)), _bf(rng, (cout, 3, 3, cin), 0.2)
    bias = rng.normal(size=cout).astype(np.float32)
    print(f"conv forward {shape}", flush=True)
    z, act = ctx.conv_forward(x, wt, bias, s)
    v = no.conv_forward(no.from_bits(x), no.from_bits(wt), bias.astype(np.float64), s)
    _close_bf16(z, v, "z")
    _close_bf16(act, no.gelu(no.from_bits(z)), "gelu(z)")  # the activation is gelu of the stored z


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_backward_data(ctx, shape):
    b, h, w, cin, cout, s = shape
    rng = np.random.default_rng(100 + sum(shape))
    dz, wt = _bf(rng, (b, h // s, w // s, cout)), _bf(rng, (cout, 3, 3, cin), 0.2)
    print(f"conv backward data {shape}", flush=True)
    dx = ctx.conv_backward_data(dz, wt, (b, h, w, cin), s)
    _close_bf16(dx, no.conv_backward_data(no.from_bits(dz), no.from_bits(wt), (b, h, w, cin), s), "dx")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("splits", [1, 5])
def test_conv_backward_weight(ctx, shape, splits):
    b, h, w, cin, cout, s = shape
    rng = np.random.default_rng(200 + sum(shape))
    x, dz = _bf(rng, (b, h, w, cin)), _bf(rng, (b, h // s, w // s, cout))
    print(f"conv backward weight {shape} splits {splits}", flush=True)
    dw, db = ctx.conv_backward_weight(x, dz, s, splits)
    want_w, want_b = no.conv_backward_weight(no.from_bits(x), no.from_bits(dz), s)
    np.testing.assert_allclose(dw, want_w, rtol=1e-4, atol=1e-4 * np.abs(want_w).max())
    np.testing.assert_allclose(db, want_b, rtol=1e-4, atol=1e-4 * np.abs(want_b).max())
    dw2, db2 = ctx.conv_backward_weight(x, dz, s, splits)
    assert np.array_equal(dw, dw2) and np.array_equal(db, db2)  # fixed reduction order


def _device_batch(ctx, arrays):
    ptrs = []
    for arr in arrays:
        p = ctx.alloc(arr.nbytes)
        ctx.upload(p, np.ascontiguousarray(arr, dtype=np.float64))
        ptrs.append(p)
    return ptrs


def _params(seed=3, sub=1):
    rng = np.random.default_rng(seed)
    p = no.init_params(seed, sub)
    lay, _ = no.layout(sub)
    for w, b, shape in lay:  # nonzero biases so every path carries signal
        p[b:b + shape[0]] = rng.normal(size=shape[0]) * 0.05
    return p


def _target(a):
    return 0.012 * no.position(a.shape[1])[None] * (6.0 / a) ** 0.3


# The full-resolution network at n = 64 and the half-resolution one (sub = 2:
# 2 x 2 blocks of cells per network pixel, a four-channel sub-pixel head) at
# n = 128, whose levels then have the same sides.
@pytest.fixture(scope="module", params=[1, 2], ids=["full-res", "half-res"])
def net_case(ctx, request):
    sub = request.param
    n, seeds = 64 * sub, [5, 6]
    a = np.stack([two_phase(n, seed=s) for s in seeds])
    u = _target(a)
    a_ptrs, u_ptrs = _device_batch(ctx, a), _device_batch(ctx, u)
    yield n, sub, a, u, a_ptrs, u_ptrs
    for p in a_ptrs + u_ptrs:
        ctx.free(p)


def test_net_gradient_matches_oracle(ctx, net_case):
    n, sub, a, u, a_ptrs, u_ptrs = net_case
    params = _params(3, sub)
    net = capi.Net(ctx, n, max_batch=4, train=True, sub=sub)
    net.set_params(params)
    print(f"net gradient n={n} sub={sub} batch 2", flush=True)
    loss, g = net.gradient(a_ptrs, u_ptrs)
    net.close()
    loss_e, g_e = no.Net(params, emulate=True).gradient(a, u)
    loss_x, g_x = no.Net(params).gradient(a, u)
    # On this case FP32 instead of FP64 accumulation in the emulated oracle
    # alone moves the loss by 1e-4 (relative) and the per-layer gradients by
    # up to 7e-3 (bf16 rounding flips); emulated and exact FP64 gradients
    # differ by up to 1.3e-2.  Bounds: about 3x those.
    assert abs(loss - loss_e) <= 5e-4 * loss_e
    lay, _ = no.layout(sub)
    for i, (w, b, shape) in enumerate(lay):
        sl = slice(w, b + shape[0])
        assert rel_l2(g[sl], g_e[sl]) < 2e-2, f"layer {i} vs emulated oracle"
        assert rel_l2(g[sl], g_x[sl]) < 5e-2, f"layer {i} vs exact oracle"


def test_net_predict_matches_oracle(ctx, net_case):
    n, sub, a, u, a_ptrs, _ = net_case
    params = _params(4, sub)
    net = capi.Net(ctx, n, max_batch=1, sub=sub)  # two chunks of one
    net.set_params(params)
    out = [ctx.alloc(n * n * 8) for _ in a_ptrs]
    f = [1.0, 2.5]
    print(f"net predict n={n} sub={sub}", flush=True)
    net.predict_device(a_ptrs, f, out)
    got = np.stack([np.empty((n, n)) for _ in out])
    for i, p in enumerate(out):
        ctx.download(got[i], p)
        ctx.free(p)
    net.close()
    want = no.Net(params, emulate=True).predict(a, np.array(f))
    exact = no.Net(params).predict(a, np.array(f))
    # The random-init output is a small residue of large cancelling sums: on
    # this case FP32 instead of FP64 accumulation alone moves the emulated
    # prediction by 5.6e-3 (bf16 rounding flips), and the emulated prediction
    # differs from exact FP64 by 9.5e-3.  Indexing mistakes give O(1).
    assert rel_l2(got, want) < 2e-2
    assert rel_l2(got, exact) < 4e-2
    np.testing.assert_array_equal(net_param_roundtrip(ctx, n, params, sub), params)


def net_param_roundtrip(ctx, n, params, sub):
    net = capi.Net(ctx, n, max_batch=1, sub=sub)
    net.set_params(params)
    out = net.get_params()
    net.close()
    return out


def test_adam_steps_reduce_loss(ctx, net_case):
    n, sub, a, u, a_ptrs, u_ptrs = net_case
    net = capi.Net(ctx, n, max_batch=2, train=True, sub=sub)
    net.set_params(_params(5, sub))
    print(f"net train steps n={n} sub={sub}", flush=True)
    losses = [net.train_step(a_ptrs, u_ptrs, 1e-3) for _ in range(30)]
    net.close()
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]


def test_net_rejects_bad_shapes(ctx):
    with pytest.raises(ValueError, match="multiple of 64"):
        capi.Net(ctx, 96, 1)
    with pytest.raises(ValueError, match="multiple of 128"):
        capi.Net(ctx, 64, 1, sub=2)
    net = capi.Net(ctx, 64, 2)
    with pytest.raises(ValueError, match="without training buffers"):
        net.train_step([0], [0], 1e-3)
    with pytest.raises(ValueError, match="wrong length"):
        net.set_params(np.zeros(10, dtype=np.float32))
    net.close()
    with pytest.raises(ValueError, match="conv shape"):
        ctx.conv_forward(np.zeros((1, 8, 8, 12), np.uint16), np.zeros((8, 3, 3, 12), np.uint16),
                         np.zeros(8, np.float32), 1)
