"""Learned warm start: a BF16 conv encoder-decoder that predicts u0.

The network (C ABI ``darcy_net_*`` in include/darcy_hip.h) maps the
permeability of an n x n problem to an initial guess u0.  With sub-pixel
factor 1 it runs on the grid itself (n a multiple of 64); with factor 2 at
half the grid's resolution (n a multiple of 128), each network pixel taking
a 2 x 2 block's four ln a values in; its output is then a bilinearly
interpolated channel plus a detail channel per cell of the block; CG then runs to the requested tolerance from u0, so the solution is the
solver's, and a poor guess only costs iterations.  It has no counterpart in
the reference, whose solve starts from u = 0 (solver.hpp:99).

Training pairs come from this solver: two-phase (12/3) GRF media from seeds
0-19999 at 256 x 256, solved in FP64 to 1e-10.  Training runs on the GPU with
the library's own forward and backward kernels and Adam.  Evaluation uses the
held-out seeds 20000-20999.

    python -m darcy.learned train [--epochs E] [--out weights.npz]
    python -m darcy.learned eval  [--weights weights.npz]

Weights are stored as a plain ``.npz``: the conv weights as the bf16 values
the GPU computes with, biases and the head in FP32, the sub-pixel factor
(``sub``, 1 when absent) and training metadata; loaded with
``allow_pickle=False``.
"""

from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time
from typing import Sequence

import numpy as np

from . import capi
from .fields import gaussian_random_field, two_phase_permeability

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_256.npz")
TRAIN_SEEDS = (0, 20000)
EVAL_SEEDS = (20000, 21000)

# Channels of the U-Net levels and (cin, cout) of its convolutions in
# parameter order: D_0..D_6, U_5..U_0, the head (1 output, or 5 at half
# resolution: the interpolated channel and the four cells' detail).
_CHAN = [32, 64, 128, 256, 256, 256, 256]


def _convs(sub: int) -> list[tuple[int, int]]:
    return ([(8 if l == 0 else _CHAN[l - 1], _CHAN[l]) for l in range(len(_CHAN))]
            + [(_CHAN[i + 1], _CHAN[i]) for i in range(len(_CHAN) - 2, -1, -1)] + [(32, 1 if sub == 1 else 5)])


def param_count(sub: int = 1) -> int:
    return sum(cout * 9 * cin + cout for cin, cout in _convs(sub))


def network_flops(n: int, sub: int = 1) -> float:
    """FLOPs of one n x n prediction: 2 * pixels * 9 * cin * cout summed over
    the convolutions (the first conv's 8 staged input channels counted) and
    the head."""
    m = n // sub
    total = 0.0
    for l in range(len(_CHAN)):  # D_l produces level l
        total += 2.0 * (m >> l) ** 2 * 9 * (8 if l == 0 else _CHAN[l - 1]) * _CHAN[l]
    for i in range(len(_CHAN) - 2, -1, -1):  # U_i runs at level i + 1
        total += 2.0 * (m >> (i + 1)) ** 2 * 9 * _CHAN[i + 1] * _CHAN[i]
    return total + 2.0 * m * m * 9 * 32 * (1 if sub == 1 else 5)


def sub_of_count(count: int) -> int:
    """The sub-pixel factor whose network has `count` parameters."""
    for sub in (1, 2):
        if count == param_count(sub):
            return sub
    raise ValueError(f"{count} parameters fit no warm-start network")


def init_params(seed: int, sub: int = 1) -> np.ndarray:
    """He-normal weights for GELU, zero biases; the first conv's unused input
    channels (sub^2 + 2 .. 7) are zero; the head starts small, its detail
    channels at zero (a smooth first output)."""
    rng = np.random.default_rng(seed)
    out = []
    convs = _convs(sub)
    used = sub * sub + 2
    for i, (cin, cout) in enumerate(convs):
        fan_in = 9 * (used if i == 0 else cin)
        std = math.sqrt(2.0 / fan_in) if i < len(convs) - 1 else 0.1 * math.sqrt(1.0 / fan_in)
        w = rng.normal(0.0, std, size=(cout, 3, 3, cin)).astype(np.float32)
        if i == 0:
            w[..., used:] = 0.0
        if i == len(convs) - 1:
            w[1:] = 0.0
        out += [w.ravel(), np.zeros(cout, dtype=np.float32)]
    p = np.concatenate(out)
    assert p.size == param_count(sub)
    return p


def _medium(args: tuple[int, int]) -> np.ndarray:
    n, seed = args
    return two_phase_permeability(gaussian_random_field(n, seed=seed))


def media(n: int, seeds: Sequence[int], workers: int = 16) -> np.ndarray:
    """Two-phase GRF media for `seeds`, generated in worker processes."""
    seeds = list(seeds)
    if workers <= 1 or len(seeds) < 64:
        return np.stack([_medium((n, s)) for s in seeds])
    with mp.get_context("spawn").Pool(workers) as pool:
        return np.stack(pool.map(_medium, [(n, s) for s in seeds], chunksize=64))


def _fp32_index(sub: int = 1) -> np.ndarray:
    """Parameters the GPU uses in FP32: every bias and the head's weights."""
    idx, o = [], 0
    convs = _convs(sub)
    for i, (cin, cout) in enumerate(convs):
        nw = cout * 9 * cin
        if i == len(convs) - 1:
            idx.append(np.arange(o, o + nw))
        idx.append(np.arange(o + nw, o + nw + cout))
        o += nw + cout
    return np.concatenate(idx).astype(np.int64)


def _bf16_bits(x: np.ndarray) -> np.ndarray:
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def save_params(path: str, params: np.ndarray, **meta) -> None:
    """The conv weights as the bf16 the GPU computes with; biases and the head
    in FP32; the sub-pixel factor follows from the parameter count."""
    sub = sub_of_count(np.asarray(params).size)
    idx = _fp32_index(sub)
    np.savez(path, weights_bf16=_bf16_bits(params), fp32_index=idx,
             fp32_values=np.asarray(params, dtype=np.float32)[idx], sub=np.int64(sub), **meta)


def load_weights(path: str = DEFAULT_WEIGHTS) -> tuple[np.ndarray, int]:
    """(parameters, sub-pixel factor) of a weight file."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: no trained warm-start weights (python -m darcy.learned train)")
    with np.load(path, allow_pickle=False) as z:
        if "params" in z:
            p = np.asarray(z["params"], dtype=np.float32)
        else:
            p = (z["weights_bf16"].astype(np.uint32) << 16).view(np.float32).copy()
            p[z["fp32_index"]] = z["fp32_values"]
        sub = int(z["sub"]) if "sub" in z else 1
    if p.size != param_count(sub):
        raise ValueError(f"{path}: {p.size} parameters, expected {param_count(sub)}")
    return p, sub


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    return load_weights(path)[0]


class WarmStart:
    """The trained network on one device context, for n x n problems."""

    def __init__(self, ctx: capi.Context, n: int = 256, max_batch: int = 1000, params: np.ndarray | None = None,
                 weights: str = DEFAULT_WEIGHTS):
        self.ctx, self.n = ctx, n
        if params is None:
            params, self.sub = load_weights(weights)
        else:
            self.sub = sub_of_count(np.asarray(params).size)
        self.net = capi.Net(ctx, n, max_batch, sub=self.sub)
        self.net.set_params(params)

    def close(self) -> None:
        self.net.close()

    def predict_device(self, a_ptrs: Sequence[int], u0_ptrs: Sequence[int], f: Sequence[float] | float = 1.0) -> None:
        """u0 (device) from a (device), for n x n problems."""
        fs = [float(f)] * len(a_ptrs) if np.isscalar(f) else [float(x) for x in f]
        self.net.predict_device(a_ptrs, fs, u0_ptrs)

    def predict(self, a: np.ndarray, f: float = 1.0) -> np.ndarray:
        """u0 for host permeabilities of shape (count, n, n) (or (n, n))."""
        a = np.asarray(a, dtype=np.float64)
        single = a.ndim == 2
        a = a[None] if single else a
        count, cells = a.shape[0], self.n * self.n
        if a.shape[1:] != (self.n, self.n):
            raise ValueError(f"expected ({self.n}, {self.n}) media, got {a.shape[1:]}")
        base_a, base_u = self.ctx.alloc(a.nbytes), self.ctx.alloc(a.nbytes)
        try:
            self.ctx.upload(base_a, a)
            self.predict_device([base_a + 8 * cells * i for i in range(count)],
                                [base_u + 8 * cells * i for i in range(count)], f)
            out = np.empty_like(a)
            self.ctx.download(out, base_u)
        finally:
            self.ctx.free(base_a)
            self.ctx.free(base_u)
        return out[0] if single else out


# ------------------------------------------------------------ training

class DeviceSet:
    """Media and FP64 solutions resident on the device, one slab each."""

    def __init__(self, ctx: capi.Context, a: np.ndarray, precision: str = "fp64"):
        self.ctx, self.count, self.n = ctx, a.shape[0], a.shape[1]
        cells = self.n * self.n
        self.a_base, self.u_base = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
        ctx.upload(self.a_base, a)
        self.a = [self.a_base + 8 * cells * i for i in range(self.count)]
        self.u = [self.u_base + 8 * cells * i for i in range(self.count)]
        res = ctx.solve_batch_device([self.n] * self.count, self.a, self.u, precision=precision)
        if not res["converged"].all():
            raise RuntimeError("training solves did not converge")
        self.iterations = res["iterations"]

    def free(self) -> None:
        self.ctx.free(self.a_base)
        self.ctx.free(self.u_base)


def train(ctx: capi.Context, data: DeviceSet, epochs: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
          log=print, sub: int = 1, spike: float = 10.0) -> tuple[np.ndarray, list]:
    """Adam with linear warm-up and cosine decay; returns (params, history).

    Rollback: the parameters at the end of each epoch whose mean loss is the
    lowest so far are kept; an epoch whose mean loss exceeds `spike` times
    that (or is not finite) is undone -- the kept parameters are restored,
    Adam's moments restart from zero and the rate is halved for the rest of
    the run.  Training under the energy-norm loss otherwise collapses now
    and then after a loss spike to a near-zero output it never leaves (the
    first epoch of such a collapse sits 80-500 times above the best;
    spikes the run recovers from by itself stay below 5 times)."""
    net = capi.Net(ctx, data.n, batch, train=True, sub=sub)
    net.set_params(init_params(seed, sub))
    rng = np.random.default_rng(seed)
    steps_per_epoch = data.count // batch
    total, warm = epochs * steps_per_epoch, min(500, epochs * steps_per_epoch // 10)
    history, step, t0 = [], 0, time.perf_counter()
    scale, best_loss, best_params = 1.0, math.inf, None
    for ep in range(epochs):
        order = rng.permutation(data.count)
        losses = []
        for s in range(steps_per_epoch):
            idx = order[s * batch:(s + 1) * batch]
            rate = scale * lr * min(1.0, (step + 1) / max(1, warm)) * 0.5 * (1.0 + math.cos(math.pi * step / total))
            want = s % 50 == 0 or s == steps_per_epoch - 1
            loss = net.train_step([data.a[i] for i in idx], [data.u[i] for i in idx], rate, want_loss=want)
            if want:
                losses.append(loss)
            step += 1
        mean = float(np.mean(losses))
        entry = {"epoch": ep + 1, "loss": mean, "seconds": time.perf_counter() - t0}
        if best_params is not None and not (mean <= spike * best_loss):
            net.set_params(best_params)  # also zeroes Adam's moments
            scale *= 0.5
            entry["rolled_back"] = True
            entry["rate_scale"] = scale
        elif mean < best_loss:
            best_loss, best_params = mean, net.get_params()
        history.append(entry)
        log(json.dumps(history[-1]))
    params = net.get_params()
    net.close()
    return params, history


def evaluate(ctx: capi.Context, params: np.ndarray, a: np.ndarray, precision: str = "mixed",
             cycle_index: int = 2) -> dict:
    """Iterations from u0 = 0 and from the prediction, and the guess's error."""
    n, count, cells = a.shape[1], a.shape[0], a.shape[1] * a.shape[1]
    base_a, base_u, base_0 = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    try:
        ctx.upload(base_a, a)
        ap = [base_a + 8 * cells * i for i in range(count)]
        up = [base_u + 8 * cells * i for i in range(count)]
        u0p = [base_0 + 8 * cells * i for i in range(count)]
        ws = WarmStart(ctx, n, params=params)
        ws.predict_device(ap, u0p)
        ws.close()
        cold = {p: ctx.solve_batch_device([n] * count, ap, up, precision=p, cycle_index=cycle_index)
                for p in ("fp64", "mixed")}
        u = np.empty_like(a)
        ctx.download(u, base_u)  # the mixed solve from 0
        u0 = np.empty_like(a)
        ctx.download(u0, base_0)
        warm = ctx.solve_batch_device([n] * count, ap, up, u0_ptrs=u0p, precision=precision,
                                      cycle_index=cycle_index)
    finally:
        for p in (base_a, base_u, base_0):
            ctx.free(p)
    err = np.linalg.norm((u0 - u).reshape(count, -1), axis=1) / np.linalg.norm(u.reshape(count, -1), axis=1)
    return {
        "problems": count,
        "iterations_fp64_cold": float(cold["fp64"]["iterations"].mean()),
        "iterations_mixed_cold": float(cold["mixed"]["iterations"].mean()),
        f"iterations_{precision}_warm": float(warm["iterations"].mean()),
        "iterations_saved_vs_mixed_cold": float(cold["mixed"]["iterations"].mean() - warm["iterations"].mean()),
        "u0_rel_l2_mean": float(err.mean()),
        "u0_rel_l2_max": float(err.max()),
        "all_converged": bool(warm["converged"].all()),
    }


def main(argv: Sequence[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m darcy.learned")
    ap.add_argument("command", choices=["train", "eval"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--train-count", type=int, default=TRAIN_SEEDS[1] - TRAIN_SEEDS[0])
    ap.add_argument("--eval-count", type=int, default=EVAL_SEEDS[1] - EVAL_SEEDS[0])
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=4e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sub", type=int, choices=(1, 2), default=2,
                    help="sub-pixel factor: 2 runs the network at half the grid's resolution")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    a_eval = media(args.n, range(EVAL_SEEDS[0], EVAL_SEEDS[0] + args.eval_count), args.workers)
    a_train = None
    if args.command == "train":
        a_train = media(args.n, range(TRAIN_SEEDS[0], TRAIN_SEEDS[0] + args.train_count), args.workers)
    print(json.dumps({"media_seconds": round(time.perf_counter() - t0, 1)}), flush=True)

    with capi.Context() as ctx:
        if args.command == "train":
            t1 = time.perf_counter()
            data = DeviceSet(ctx, a_train)
            del a_train
            print(json.dumps({"pairs": data.count, "solve_seconds": round(time.perf_counter() - t1, 1),
                              "mean_iterations": float(data.iterations.mean())}), flush=True)
            params, history = train(ctx, data, args.epochs, args.batch, args.lr, args.seed,
                                    log=lambda s: print(s, flush=True), sub=args.sub)
            data.free()
        else:
            params, history = load_params(args.weights), []
        result = evaluate(ctx, params, a_eval)
        print(json.dumps(result), flush=True)
        if args.command == "train":
            os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
            save_params(args.out, params, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
                        train_seeds=np.array([TRAIN_SEEDS[0], TRAIN_SEEDS[0] + args.train_count]),
                        loss=np.array([h["loss"] for h in history]), eval_json=json.dumps(result))
            np.save(os.path.splitext(args.out)[0] + "_fp32.npy", params)
            print(json.dumps({"saved": args.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
