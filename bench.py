#!/usr/bin/env python3
"""Benchmark of the darcy solve path on one MI355X.

Workload (the headline metric's configuration): 1000 independent 256 x 256
two-phase media (GRF alpha=2, tau=3 thresholded to 12 / 3, held-out seeds
20000-20999), f = 1, solved to a 1e-10 relative residual with the
multigrid-preconditioned CG, W-cycle.  One "step" solves the whole batch from
HBM-resident permeabilities to HBM-resident solutions.

Three GPU variants are measured on the same batch:
  * "learned": the warm-start network predicts u0 (BF16 MFMA convolutions),
    then the FP32-cycle / FP64-flexible-CG solve runs from it  -> `value`
  * "mixed":   the same solve from u0 = 0                      -> `mixed`
  * "fp64":    plain FP64 MG-CG from u0 = 0 (the reference's arithmetic) -> `fp64`
and the reference CPU path (the reference's own solver headers built into
oracle/_ref, OpenMP across samples) on a bounded sample of the same media
-> `cpu_baseline`.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch 1000]

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "darcy-cpp_amd"))

from darcy.fields import gaussian_random_field, two_phase_permeability  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
# Bytes between the L2 caches and the memory fabric (an upper bound on HBM
# bytes: Infinity Cache hits are included) moved by k_batch<float> in one
# launch of the default workload's learned line (1000 x 256^2, W-cycle,
# committed weights): rocprofv3 FETCH_SIZE (x 2) + WRITE_SIZE (x 1), the
# factors that the same counters give for the library's copy kernel of 1 GiB.
# Reported for that configuration only.
TRAFFIC_DEFAULT_BYTES = 406.4e9
SEED0 = 20000

# Algorithmic bytes per fine cell (counting rule: each array a kernel must
# read or write counted once at its storage size; the operator as two face
# arrays per level; halo re-reads and scratch not counted).  The FP32-cycle
# kernel applies u += alpha p in the next direction pass, which reads p
# anyway: 8 bytes fewer than the FP64 layout of the same passes.  The
# smoothers form inv_diag from the faces they read, so the cycle reads no
# inv_diag array (2 reads per visit fewer: 2 x 2 s for W, 4/3 x 2 s for V).
B_ITER = {("fp64", 2): 292.0, ("fp64", 1): 226.7, ("mixed", 2): 206.0, ("mixed", 1): 173.3}
B_SETUP = {"fp64": 51.0, "mixed": 62.0}


def media(n: int, count: int, seed0: int) -> np.ndarray:
    out = np.empty((count, n, n))
    for i in range(count):
        out[i] = two_phase_permeability(gaussian_random_field(n, seed=seed0 + i))
    return out


def algorithmic_bytes(iters: np.ndarray, n: int, precision: str, cycle: int) -> float:
    cells = float(n * n)
    return float(iters.sum()) * cells * B_ITER[(precision, cycle)] + len(iters) * cells * B_SETUP[precision]


def run_gpu(ctx, a_ptrs, u_ptrs, n, precision, steps, warmup, cycle, warm=None, u0_ptrs=None):
    """Time `steps` batch solves; with `warm` (a WarmStart) each step first
    predicts u0 into u0_ptrs.  Returns (result, wall ms/step, solve-kernel ms,
    inference ms)."""
    sizes = [n] * len(a_ptrs)

    def step():
        inf = 0.0
        if warm is not None:
            warm.predict_device(a_ptrs, u0_ptrs)
            inf = ctx.last_kernel_ms()
        r = ctx.solve_batch_device(sizes, a_ptrs, u_ptrs, u0_ptrs=u0_ptrs if warm is not None else None,
                                   precision=precision, cycle_index=cycle)
        return r, ctx.last_kernel_ms(), inf

    res = None
    for _ in range(warmup):
        res = step()[0]
    kernel_ms, inf_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        res, k, i = step()
        kernel_ms.append(k)
        inf_ms.append(i)
    wall = time.perf_counter() - t0
    return res, wall / steps * 1e3, float(np.mean(kernel_ms)), float(np.mean(inf_ms))


def host_cores() -> tuple[int, int, str]:
    """(threads to use, cores visible, CPU model).  Threads: every core this
    process may run on, bounded by OMP_NUM_THREADS where it is set -- on the
    GPU boxes it is the CPU share that comes with one GPU (16), which the
    box's rules say pools must not exceed."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        visible = os.cpu_count() or 1
    threads = visible
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:  # pragma: no cover
        pass
    return max(1, threads), visible, model


def cpu_baseline(n: int, sample: int, cycle: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads, visible, model = host_cores()
    a = media(n, sample, SEED0)
    if pyoracle.reference_available():
        solver, kind = pyoracle.Reference(), "reference"
    else:
        solver, kind = pyoracle.Oracle(), "port"
    solver.solve_batch(a[: min(4, sample)], threads=threads, cycle_index=cycle)  # warm-up
    t0 = time.perf_counter()
    _, iters, _, conv = solver.solve_batch(a, threads=threads, cycle_index=cycle)
    dt = time.perf_counter() - t0
    return {
        "value": sample / dt,
        "unit": "solves/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{sample} of the {n}x{n} media (seeds {SEED0}-{SEED0 + sample - 1}), FP64 W-cycle MG-CG, "
                  f"OpenMP dynamic across samples on {threads} threads; mean {iters.mean():.1f} iterations, "
                  f"{dt:.1f} s wall",
        "cores_visible": visible,
        "cpu_model": model,
    }


def suite_large(args) -> int:
    """Large single solves (1024^2 and 4096^2, two-phase 12/3 and 1e3/1e-3
    media, seed 0): FP64 and FP32-cycle paths, event time, iterations and,
    with --oracle, the relative L2 distance to the CPU oracle's solution."""
    oracle = None
    if args.oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        oracle = pyoracle.Oracle()
    ctx = _context(args)
    if args.no_graph:
        ctx.set_tuning("grid_graph", 0)
    for n in args.sizes:
        for medium, hi, lo in (("two_phase_12_3", 12.0, 3.0), ("contrast_1e6", 1e3, 1e-3)):
            if args.media and medium not in args.media:
                continue
            a = two_phase_permeability(gaussian_random_field(n, seed=0), hi=hi, lo=lo)
            ab, ub = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
            ctx.upload(ab, a)
            ref = oracle.solve(a, max_iter=20000) if oracle is not None else None
            for precision in args.precisions:
                ctx.solve_device(n, ab, ub, precision=precision, max_iter=20000)  # warm-up
                times = []
                for _ in range(max(1, args.steps)):
                    t0 = time.perf_counter()
                    r = ctx.solve_device(n, ab, ub, precision=precision, max_iter=20000)
                    times.append((time.perf_counter() - t0) * 1e3)
                u = np.empty_like(a)
                ctx.download(u, ub)
                line = {"suite": "large", "n": n, "medium": medium, "precision": precision,
                        "iterations": r["iterations"], "levels": r["levels"], "converged": r["converged"],
                        "residual": r["residual"], "ms_wall_min": round(min(times), 3),
                        "ms_event": round(ctx.last_kernel_ms(), 3)}
                if ref is not None:
                    line["oracle_iterations"] = ref.iterations
                    line["rel_l2_vs_oracle"] = float(np.linalg.norm(u - ref.u) / np.linalg.norm(ref.u))
                print(json.dumps(line), flush=True)
            ctx.free(ab)
            ctx.free(ub)
    ctx.close()
    return 0


def suite_phases(args) -> int:
    """Where an iteration of the persistent kernel goes: the default batch
    solved for a fixed number of iterations (tol far below reach), with the
    multigrid cycle and with no preconditioner (z = r, one extra copy pass),
    in each precision; (time at 21 iterations - time at 1) / 20 is the cost
    of one iteration, kernel event time, best of 3."""
    n, batch = args.n, args.batch
    a = media(n, batch, SEED0)
    ctx = _context(args)
    cells = n * n
    ab, ub = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    ctx.upload(ab, a)
    ap = [ab + 8 * cells * i for i in range(batch)]
    up = [ub + 8 * cells * i for i in range(batch)]
    out = {}
    for precision in ("mixed", "fp64"):
        for pc in ("multigrid", "none"):
            ms = {}
            for it in (1, 21):
                best = None
                for _ in range(3):
                    ctx.solve_batch_device([n] * batch, ap, up, tol=1e-300, max_iter=it, preconditioner=pc,
                                           precision=precision, cycle_index=args.cycle)
                    t = ctx.last_kernel_ms()
                    best = t if best is None else min(best, t)
                ms[it] = best
            out[f"{precision}/{pc}"] = {"ms_per_iteration": round((ms[21] - ms[1]) / 20, 3),
                                        "ms_setup_and_first": round(ms[1], 3)}
    ctx.free(ab)
    ctx.free(ub)
    ctx.close()
    print(json.dumps({"suite": "phases", "problems": batch, "n": n, **out}), flush=True)
    return 0


def uneven_sizes(count: int) -> np.ndarray:
    """Grid sides of the uneven batch: default_rng(0).choice over the seven sizes."""
    return np.random.default_rng(0).choice(np.array([64, 96, 128, 192, 256, 384, 512]), size=count)


def suite_uneven(args) -> int:
    """2000 problems, side drawn from {64,...,512} by seed 0, problem i on GRF
    seed i, two-phase 12/3, solved in one persistent launch per step."""
    count = args.batch if args.batch != 1000 else 2000
    sizes = uneven_sizes(count)
    media = [two_phase_permeability(gaussian_random_field(int(n), seed=i)) for i, n in enumerate(sizes)]
    ctx = _context(args)
    total = int(sum(int(n) * int(n) for n in sizes))
    ab, ub = ctx.alloc(total * 8), ctx.alloc(total * 8)
    a_ptrs, u_ptrs, off = [], [], 0
    for m in media:
        ctx.upload(ab + off * 8, m)
        a_ptrs.append(ab + off * 8)
        u_ptrs.append(ub + off * 8)
        off += m.size
    out = {}
    for precision in ("fp64", "mixed"):
        ctx.solve_batch_device([int(n) for n in sizes], a_ptrs, u_ptrs, precision=precision)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = ctx.solve_batch_device([int(n) for n in sizes], a_ptrs, u_ptrs, precision=precision)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        it = r["iterations"]
        out[precision] = {"solves_per_s": round(count / (ms * 1e-3), 1), "ms_per_step": round(ms, 2),
                          "kernel_ms": round(ctx.last_kernel_ms(), 2), "all_converged": bool(r["converged"].all()),
                          "iterations_by_size": {int(s): float(it[sizes == s].mean()) for s in np.unique(sizes)}}
    ctx.free(ab)
    ctx.free(ub)
    ctx.close()
    line = {"suite": "uneven", "count": count, "cells": total, "sizes": {int(s): int((sizes == s).sum())
                                                                         for s in np.unique(sizes)}, **out}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        solver = pyoracle.Reference() if pyoracle.reference_available() else pyoracle.Oracle()
        threads, visible, model = host_cores()
        t_total, solved = 0.0, 0
        for s in np.unique(sizes):  # the reference batch needs one n per call
            idx = np.nonzero(sizes == s)[0][: args.cpu_sample // 7 + 1]
            batch = np.stack([media[i] for i in idx])
            t0 = time.perf_counter()
            solver.solve_batch(batch, threads=threads)
            t_total += time.perf_counter() - t0
            solved += len(idx)
        line["cpu_baseline_sample"] = {"solves_per_s_on_sample": round(solved / t_total, 1), "cores": threads,
                                       "cores_visible": visible, "cpu_model": model, "problems": solved,
                                       "note": "per-size-group calls, sample of each group"}
    print(json.dumps(line), flush=True)
    return 0


def _context(args):
    """A context with the tuning table applied, then each --tune KEY=VALUE."""
    from darcy.capi import Context

    ctx = Context(0)
    for item in args.tune:
        key, _, value = item.partition("=")
        ctx.set_tuning(key, int(value))
    return ctx


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--suite", choices=["metric", "large", "uneven", "phases"], default="metric")
    ap.add_argument("--sizes", type=int, nargs="+", default=[1024, 4096])
    ap.add_argument("--oracle", action="store_true", help="large suite: compare with the CPU oracle")
    ap.add_argument("--media", nargs="+", choices=["two_phase_12_3", "contrast_1e6"], default=None,
                    help="large suite: only these media (default both)")
    ap.add_argument("--precisions", nargs="+", choices=["fp64", "mixed"], default=["fp64", "mixed"],
                    help="large suite: precisions to run")
    ap.add_argument("--no-graph", action="store_true",
                    help="large suite: launch the cycle's kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1000)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cycle", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=320)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--weights", default=None, help="warm-start weight file (default: the committed weights)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a tuning key after the tuning table (repeatable)")
    ap.add_argument("--copy-only", type=int, default=0,
                    help="only run the streaming copy kernel on this many MiB (profiler calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: independent solves shard across GPUs with no communication, "
              "so this build runs and reports N=1 only", file=sys.stderr)
        return 2

    from darcy.capi import Context

    if args.suite == "large":
        return suite_large(args)
    if args.suite == "uneven":
        return suite_uneven(args)
    if args.suite == "phases":
        return suite_phases(args)
    if args.copy_only:
        with Context(0) as c:
            print(json.dumps({"copy_MiB": args.copy_only, "GBs": c.bandwidth_probe(args.copy_only << 20, 5)}))
        return 0

    n, batch = args.n, args.batch
    a = media(n, batch, SEED0)
    ctx = _context(args)
    copy_gbs = ctx.bandwidth_probe(1 << 30, 10)
    cells = n * n
    a_base = ctx.alloc(a.nbytes)
    u_base = ctx.alloc(a.nbytes)
    ctx.upload(a_base, a)
    a_ptrs = [a_base + i * cells * 8 for i in range(batch)]
    u_ptrs = [u_base + i * cells * 8 for i in range(batch)]

    from darcy.learned import DEFAULT_WEIGHTS, WarmStart, network_flops, param_count

    warm = WarmStart(ctx, n, weights=args.weights or DEFAULT_WEIGHTS)  # fails loudly without trained weights
    net_sub = warm.sub
    u0_base = ctx.alloc(a.nbytes)
    u0_ptrs = [u0_base + i * cells * 8 for i in range(batch)]

    out = {}
    for name, precision, ws in (("fp64", "fp64", None), ("mixed", "mixed", None), ("learned", "mixed", warm)):
        res, ms_wall, ms_kernel, ms_inf = run_gpu(ctx, a_ptrs, u_ptrs, n, precision, args.steps, args.warmup,
                                                  args.cycle, ws, u0_ptrs)
        if not res["converged"].all():
            print(f"{name}: {int((~res['converged']).sum())} problems did not converge", file=sys.stderr)
            return 1
        nbytes = algorithmic_bytes(res["iterations"], n, precision, args.cycle)
        out[name] = {
            "ms_per_step": ms_wall,
            "kernel_ms": ms_kernel,
            "inference_ms": ms_inf,
            "solves_per_s": batch / (ms_wall * 1e-3),
            "mean_iterations": float(res["iterations"].mean()),
            "max_iterations": int(res["iterations"].max()),
            "alg_bytes": nbytes,
            "achieved_GBs": nbytes / (ms_kernel * 1e-3) / 1e9,
        }
    warm.close()
    ctx.free(u0_base)
    ctx.free(a_base)
    ctx.free(u_base)
    info = ctx.info()
    ctx.close()

    learned, mixed = out["learned"], out["mixed"]
    flops_per_sample = network_flops(n, net_sub)
    net_flops = flops_per_sample * batch
    line = {
        "metric": "solves/s to 1e-10 resid, 256² batch: learned+FP32-MG vs plain FP64 MG-CG",
        "value": round(learned["solves_per_s"], 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(learned["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cycle / f64 CG; bf16 network",
        "data": "synthetic: GRF two-phase media generated from seeds; network trained on seeds 0-19999",
        "config": {
            "workload": f"{batch} x {n}x{n} two-phase (12/3) GRF media, seeds {SEED0}-{SEED0 + batch - 1}, f=1, "
                        f"tol 1e-10, {'W' if args.cycle == 2 else 'V'}-cycle MG-CG; learned warm start (BF16 "
                        f"encoder-decoder, {param_count(net_sub)} parameters, "
                        f"{'full' if net_sub == 1 else 'half'}-resolution body) then FP32 cycle in FP64 "
                        f"flexible CG; network inference inside the timed region",
            "global_batch": batch,
            "grid": n,
            "parallelism": "single GPU",
        },
        "mean_iterations": learned["mean_iterations"],
        "inference_ms": round(learned["inference_ms"], 3),
        "solve_ms": round(learned["kernel_ms"], 3),
        "speedup_vs_fp64": round(learned["solves_per_s"] / out["fp64"]["solves_per_s"], 3),
        # the learned line against the FP32 cycle from u0 = 0 (below 1: the
        # network costs more time than the iterations it saves)
        "learned_vs_mixed": round(learned["solves_per_s"] / mixed["solves_per_s"], 3),
        "fp64": {
            "value": round(out["fp64"]["solves_per_s"], 2),
            "ms_per_step": round(out["fp64"]["ms_per_step"], 3),
            "mean_iterations": out["fp64"]["mean_iterations"],
            "achieved_GBs": round(out["fp64"]["achieved_GBs"], 1),
        },
        "mixed": {
            "value": round(mixed["solves_per_s"], 2),
            "ms_per_step": round(mixed["ms_per_step"], 3),
            "mean_iterations": mixed["mean_iterations"],
            "speedup_vs_fp64": round(mixed["solves_per_s"] / out["fp64"]["solves_per_s"], 3),
        },
        "network": {
            "flops_per_sample": flops_per_sample,
            "sub_pixel_factor": net_sub,
            "inference_ms": round(learned["inference_ms"], 3),
            "TFLOPs": round(net_flops / (learned["inference_ms"] * 1e-3) / 1e12, 1),
            "frac_bf16_peak": round(net_flops / (learned["inference_ms"] * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
            "iterations_saved": round(mixed["mean_iterations"] - learned["mean_iterations"], 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "darcy_hip::k_batch<float> (persistent PCG, one workgroup per problem), learned line",
            "achieved": round(learned["achieved_GBs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(learned["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": TRAFFIC_DEFAULT_BYTES if (batch, n, args.cycle) == (1000, 256, 2) else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE, scaled on the copy kernel: bytes between "
                              "the L2 caches and the memory fabric (Infinity Cache hits included), an upper "
                              "bound on HBM bytes",
            "launch_ms": round(learned["kernel_ms"], 3),
            "alg_bytes_per_launch": learned["alg_bytes"],
        },
        "device": info,
        "copy_kernel_GBs": round(copy_gbs, 1),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_sample, args.cycle)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
