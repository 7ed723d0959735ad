"""CPU oracle of the learned warm start network, in FP64 numpy.

Test infrastructure: only tests, the smoke check and the benchmark's CPU leg
use it, as the checker of the GPU network (libdarcy_hip: darcy_net_*,
darcy_nn_*).  The reference has no learned component, so nothing here
restates reference code; the network is the one the C header describes: a
U-Net whose level l has side m / 2^l and CHAN[l] channels, m = n / sub the
side of its finest level (sub = 1: the grid; sub = 2: one network pixel per
2 x 2 block of cells),

    features: the sub^2 cells' normalised ln a, 1, the cells' mean position
    enc: D_0 conv 8->32 at m, D_l stride-2 conv from level l-1 to level l
    dec: U_i (i = L-2..0) conv at level i+1 to CHAN[i] channels, nearest x2
         upsampling, + the encoder activation of level i, GELU
    head: conv 32->1 (sub = 1), or 32->5 (sub = 2): channel 0 interpolated
          bilinearly to the cells (cell-centred, indices clamped to the
          grid), plus channel 1+2di+dj at cell (2I+di, 2J+dj);
          u0 = f * pos * y / 64

and the training loss (sum e^2 + 1000 e^T A e) / cells, e = pos y - 64 u, A the
finite-volume operator of the permeability.

Two modes share one code path:

* ``emulate=False``: exact FP64 math.  Its gradient is pinned by central
  finite differences of its own loss; its convolution by a direct
  seven-loop sum and by the adjoint identity <conv x, dz> = <x, conv^T dz>.
* ``emulate=True``: rounds to bf16 (round to nearest even) exactly where the
  GPU stores bf16 -- input features, conv weights, every stored
  pre-activation z and activation gelu(z), every activation gradient -- so
  the GPU result differs from it only by FP32 accumulation order.
"""

from __future__ import annotations

import numpy as np

CHAN = [32, 64, 128, 256, 256, 256, 256]
LEVELS = len(CHAN)
IN_C = 8
HEAD_CIN = 32
SCALE = 1.0 / 64.0
ENERGY_WEIGHT = 1000.0
LN3, LN12 = np.log(3.0), np.log(12.0)


def conv_specs() -> list[tuple[int, int, int, int]]:
    """(cin, cout, stride, input level) of D_0..D_{L-1}, U_{L-2}..U_0, in parameter order."""
    out = [(IN_C if l == 0 else CHAN[l - 1], CHAN[l], 1 if l == 0 else 2, max(l - 1, 0)) for l in range(LEVELS)]
    out += [(CHAN[i + 1], CHAN[i], 1, i + 1) for i in range(LEVELS - 2, -1, -1)]
    return out


def layout(sub: int = 1) -> tuple[list[tuple[int, int, tuple[int, ...]]], int]:
    """[(w_offset, b_offset, w_shape)] for the convs and the head, and the total."""
    out, o = [], 0
    for cin, cout, _, _ in conv_specs() + [(HEAD_CIN, 1 if sub == 1 else 5, 1, 0)]:
        shape = (cout, 3, 3, cin)
        w = o
        o += int(np.prod(shape))
        out.append((w, o, shape))
        o += cout
    return out, o


def sub_of(params: np.ndarray) -> int:
    """The sub-pixel factor whose network has params.size parameters."""
    for sub in (1, 2):
        if params.size == layout(sub)[1]:
            return sub
    raise ValueError(f"{params.size} parameters fit no network")


def unpack(params: np.ndarray) -> list[tuple[np.ndarray, np.ndarray]]:
    lay, total = layout(sub_of(params))
    assert params.shape == (total,)
    res = []
    for w, b, shape in lay:
        res.append((params[w:b].reshape(shape).astype(np.float64), params[b:b + shape[0]].astype(np.float64)))
    return res


def pack(pairs: list[tuple[np.ndarray, np.ndarray]]) -> np.ndarray:
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in pairs])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) through float32; returns float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """bf16 bit patterns (uint16) of x, rounded to nearest even."""
    return (bf16(x).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _gelu_s(z):
    """(1 + tanh u) / 2 = 1 / (1 + exp(-2u)), u = sqrt(2/pi) (z + 0.044715 z^3)."""
    c = np.sqrt(2.0 / np.pi)
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(-2.0 * c * (z + 0.044715 * z ** 3)))


def gelu(z):
    """GELU, tanh form 0.5 z (1 + tanh u), written as z s to avoid cancellation."""
    return z * _gelu_s(z)


def gelu_grad(z):
    c = np.sqrt(2.0 / np.pi)
    s = _gelu_s(z)
    return s + 2.0 * z * s * (1.0 - s) * c * (1.0 + 3.0 * 0.044715 * z * z)


def position(n: int) -> np.ndarray:
    """sin(pi (i + 1/2) / n) sin(pi (j + 1/2) / n), rounded to float32 as the GPU does."""
    s = np.sin(np.pi * (np.arange(n) + 0.5) / n)
    return np.outer(s, s).astype(np.float32).astype(np.float64)


# ------------------------------------------------------------ convolutions

def _patches(x: np.ndarray, stride: int) -> np.ndarray:
    b, h, w, c = x.shape
    oh, ow = h // stride, w // stride
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    p = np.empty((b, oh, ow, 3, 3, c))
    for ky in range(3):
        for kx in range(3):
            p[:, :, :, ky, kx, :] = xp[:, ky:ky + stride * oh:stride, kx:kx + stride * ow:stride, :]
    return p


def conv_forward(x: np.ndarray, w: np.ndarray, bias: np.ndarray, stride: int) -> np.ndarray:
    """z[b, oy, ox, co] = bias[co] + sum w[co, ky, kx, ci] x[b, s oy + ky - 1, s ox + kx - 1, ci] (zero outside)."""
    cout = w.shape[0]
    p = _patches(x, stride)
    z = p.reshape(-1, p.shape[3] * p.shape[4] * p.shape[5]) @ w.reshape(cout, -1).T + bias
    return z.reshape(p.shape[0], p.shape[1], p.shape[2], cout)


def conv_backward_data(dz: np.ndarray, w: np.ndarray, in_shape: tuple[int, ...], stride: int) -> np.ndarray:
    b, h, wd, c = in_shape
    oh, ow = dz.shape[1], dz.shape[2]
    dp = (dz.reshape(-1, dz.shape[3]) @ w.reshape(w.shape[0], -1)).reshape(b, oh, ow, 3, 3, c)
    dxp = np.zeros((b, h + 2, wd + 2, c))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + stride * oh:stride, kx:kx + stride * ow:stride, :] += dp[:, :, :, ky, kx, :]
    return dxp[:, 1:-1, 1:-1, :]


def conv_backward_weight(x: np.ndarray, dz: np.ndarray, stride: int) -> tuple[np.ndarray, np.ndarray]:
    p = _patches(x, stride)
    cout = dz.shape[3]
    dw = dz.reshape(-1, cout).T @ p.reshape(-1, p.shape[3] * p.shape[4] * p.shape[5])
    return dw.reshape(cout, 3, 3, x.shape[3]), dz.reshape(-1, cout).sum(axis=0)


def conv_direct(x: np.ndarray, w: np.ndarray, bias: np.ndarray, stride: int) -> np.ndarray:
    """The definition, one output at a time (tiny shapes only)."""
    b, h, wd, cin = x.shape
    cout = w.shape[0]
    oh, ow = h // stride, wd // stride
    z = np.zeros((b, oh, ow, cout))
    for bb in range(b):
        for oy in range(oh):
            for ox in range(ow):
                for co in range(cout):
                    s = bias[co]
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = stride * oy + ky - 1, stride * ox + kx - 1
                            if 0 <= iy < h and 0 <= ix < wd:
                                for ci in range(cin):
                                    s += w[co, ky, kx, ci] * x[bb, iy, ix, ci]
                    z[bb, oy, ox, co] = s
    return z


def operator(a: np.ndarray, x: np.ndarray) -> np.ndarray:
    """A x for a batch (B, n, n): harmonic interior faces, boundary faces
    carrying the cell's own a (the solver's operator, without the h^2 of b)."""
    hm = lambda p, q: 2.0 * p * q / (p + q)  # noqa: E731
    out = np.zeros_like(x)
    fe, fn = hm(a[:, :, :-1], a[:, :, 1:]), hm(a[:, :-1, :], a[:, 1:, :])
    dx, dy = x[:, :, :-1] - x[:, :, 1:], x[:, :-1, :] - x[:, 1:, :]
    out[:, :, :-1] += fe * dx
    out[:, :, 1:] -= fe * dx
    out[:, :-1, :] += fn * dy
    out[:, 1:, :] -= fn * dy
    out[:, :, 0] += a[:, :, 0] * x[:, :, 0]
    out[:, :, -1] += a[:, :, -1] * x[:, :, -1]
    out[:, 0, :] += a[:, 0, :] * x[:, 0, :]
    out[:, -1, :] += a[:, -1, :] * x[:, -1, :]
    return out


def upsample2(c: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(c, 2, axis=1), 2, axis=2)


def down_sum(d: np.ndarray) -> np.ndarray:
    b, h, w, c = d.shape
    return d.reshape(b, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def _bilinear_matrix(m: int) -> np.ndarray:
    """(2m, m) cell-centred linear interpolation: fine index i = 2I + d takes
    3/4 of I and 1/4 of its neighbour on side d (I - 1 or I + 1, clamped)."""
    P = np.zeros((2 * m, m))
    for i in range(2 * m):
        I = i // 2
        side = min(max(I + (1 if i % 2 else -1), 0), m - 1)
        P[i, I] += 0.75
        P[i, side] += 0.25
    return P


def sub2_out(yc: np.ndarray) -> np.ndarray:
    """(B, m, m, 5) -> (B, 2m, 2m): bilinear channel 0 plus the detail channels."""
    P = _bilinear_matrix(yc.shape[1])
    return np.einsum("ia,bac,jc->bij", P, yc[..., 0], P) + pixel_shuffle(yc[..., 1:], 2)


def sub2_adjoint(d: np.ndarray) -> np.ndarray:
    """The adjoint of sub2_out: (B, 2m, 2m) -> (B, m, m, 5)."""
    P = _bilinear_matrix(d.shape[1] // 2)
    return np.concatenate([np.einsum("ia,bij,jc->bac", P, d, P)[..., None], pixel_unshuffle(d, 2)], axis=-1)


def pixel_shuffle(y: np.ndarray, sub: int) -> np.ndarray:
    """(B, m, m, sub^2) -> (B, sub m, sub m): channel di*sub+dj to cell (sub I+di, sub J+dj)."""
    b, m, _, _ = y.shape
    return y.reshape(b, m, m, sub, sub).transpose(0, 1, 3, 2, 4).reshape(b, sub * m, sub * m)


def pixel_unshuffle(d: np.ndarray, sub: int) -> np.ndarray:
    """The inverse (and adjoint) of pixel_shuffle."""
    b, n, _ = d.shape
    m = n // sub
    return d.reshape(b, m, sub, m, sub).transpose(0, 1, 3, 2, 4).reshape(b, m, m, sub * sub)


# ------------------------------------------------------------ the network

def features(a: np.ndarray, sub: int = 1) -> np.ndarray:
    """(B, n, n) permeability -> (B, n/sub, n/sub, 8) features before rounding:
    the sub^2 cells' (ln a - mu) / sd (row-major, as float32), 1, and the
    cells' mean position (summed in row-major order in float32)."""
    bsz, n, _ = a.shape
    m = n // sub
    x = np.zeros((bsz, m, m, 8))
    mu, sd = 0.5 * (LN3 + LN12), 0.5 * (LN12 - LN3)
    la = ((np.log(a) - mu) / sd).astype(np.float32)
    pos = position(n).astype(np.float32)
    ps = np.zeros((m, m), dtype=np.float32)
    for r in range(sub * sub):
        di, dj = divmod(r, sub)
        x[..., r] = la[:, di::sub, dj::sub]
        ps = (ps + pos[di::sub, dj::sub]).astype(np.float32)
    x[..., sub * sub] = 1.0
    x[..., sub * sub + 1] = ps * np.float32(1.0 / (sub * sub))
    return x


class Net:
    """Forward, loss and backward of the warm-start network on a batch."""

    def __init__(self, params: np.ndarray, emulate: bool = False):
        self.emulate = emulate
        self.sub = sub_of(np.asarray(params))
        self.layers = unpack(np.asarray(params, dtype=np.float64))
        self.R = bf16 if emulate else (lambda v: v)
        nconv = 2 * LEVELS - 1
        # Conv weights as the GPU uses them (bf16 copies); the head stays FP32.
        self.w = [(self.R(w), b) for w, b in self.layers[:nconv]] + [self.layers[nconv]]

    @staticmethod
    def u_index(i: int) -> int:
        return 2 * LEVELS - 2 - i

    def forward(self, a: np.ndarray) -> dict:
        R, s = self.R, {}
        x = R(features(a, self.sub))
        s["x"] = x
        for l in range(LEVELS):
            w, b = self.w[l]
            z = R(conv_forward(x, w, b, 1 if l == 0 else 2))
            s[f"z{l}"], s[f"a{l}"] = z, R(gelu(z))
            x = s[f"a{l}"]
        for i in range(LEVELS - 2, -1, -1):
            w, b = self.w[self.u_index(i)]
            c = R(conv_forward(x, w, b, 1))
            z = R(upsample2(c) + s[f"a{i}"])
            s[f"zm{i}"], s[f"am{i}"] = z, R(gelu(z))
            x = s[f"am{i}"]
        w, b = self.w[-1]
        yc = conv_forward(x, w, b, 1)
        s["y"] = yc[..., 0] if self.sub == 1 else sub2_out(yc)
        return s

    def loss(self, a: np.ndarray, u: np.ndarray) -> tuple[float, dict]:
        s = self.forward(a)
        pos = position(a.shape[1])
        t = u / SCALE
        if self.emulate:
            t = t.astype(np.float32).astype(np.float64)
        e = pos * s["y"] - t
        ae = operator(a, e)
        s["dy"] = 2.0 * (e + ENERGY_WEIGHT * ae) / e.size * pos
        return float((np.sum(e * e) + ENERGY_WEIGHT * np.sum(e * ae)) / e.size), s

    def gradient(self, a: np.ndarray, u: np.ndarray) -> tuple[float, np.ndarray]:
        R = self.R
        loss, s = self.loss(a, u)
        grads: list = [None] * (2 * LEVELS)
        dy = s["dy"][..., None] if self.sub == 1 else sub2_adjoint(s["dy"])
        w, _ = self.w[-1]
        grads[-1] = conv_backward_weight(s["am0"], dy, 1)
        dm = R(conv_backward_data(dy, w, s["am0"].shape, 1))
        # Decoder, finest merge first; dzm[i] is also the skip gradient of a_i.
        dzm = {}
        for i in range(LEVELS - 1):
            c = self.u_index(i)
            dz = R(dm * gelu_grad(s[f"zm{i}"]))
            dzm[i] = dz
            dc = R(down_sum(dz))
            xin = s[f"a{LEVELS - 1}"] if i == LEVELS - 2 else s[f"am{i + 1}"]
            grads[c] = conv_backward_weight(xin, dc, 1)
            dm = R(conv_backward_data(dc, self.w[c][0], xin.shape, 1))
        # Encoder, deepest first: dm is the gradient of a_{L-1}.
        d = R(dm * gelu_grad(s[f"z{LEVELS - 1}"]))
        for l in range(LEVELS - 1, -1, -1):
            x = s[f"a{l - 1}"] if l > 0 else s["x"]
            grads[l] = conv_backward_weight(x, d, 1 if l == 0 else 2)
            if l == 0:
                break
            dx = R(conv_backward_data(d, self.w[l][0], x.shape, 2))
            dx = R(dx + dzm[l - 1])
            d = R(dx * gelu_grad(s[f"z{l - 1}"]))
        return loss, pack(grads)

    def predict(self, a: np.ndarray, f: np.ndarray | float = 1.0) -> np.ndarray:
        s = self.forward(a)
        f = np.broadcast_to(np.asarray(f, dtype=np.float64), (a.shape[0],))
        return f[:, None, None] * SCALE * (position(a.shape[1]) * s["y"])


def init_params(seed: int, sub: int = 1) -> np.ndarray:
    """He-normal conv weights for GELU, zero biases; unused input channels
    (sub^2 + 2 .. 7) zero."""
    rng = np.random.default_rng(seed)
    lay, total = layout(sub)
    used = sub * sub + 2
    p = np.zeros(total, dtype=np.float32)
    for i, (w, b, shape) in enumerate(lay):
        fan_in = 9 * (used if i == 0 else shape[3])
        std = np.sqrt(2.0 / fan_in) if i < len(lay) - 1 else 0.1 * np.sqrt(1.0 / fan_in)
        wt = rng.normal(0.0, std, size=shape).astype(np.float32)
        if i == 0:
            wt[..., used:] = 0.0
        if i == len(lay) - 1:
            wt[1:] = 0.0  # the detail channels start at zero
        p[w:b] = wt.ravel()
    return p
