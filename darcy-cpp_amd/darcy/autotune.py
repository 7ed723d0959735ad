"""Launch-shape autotuner for the solve path.

The C ABI exposes launch-shape knobs on a context (darcy_ctx_set_tuning);
none of them changes a result, only how the work is launched:

  grid_min_n      single multigrid solves with n above this use the
                  multi-workgroup grid path, the others the persistent
                  one-workgroup-per-problem kernel
  grid_tail_side  on the grid path, levels with side <= this run inside one
                  workgroup (the rest phase by phase over the whole GPU)
  grid_graph      replay the grid path's cycle from a captured HIP graph
  batch_step_cells  cells per row-pipeline step of the persistent kernel's
                  levels in global memory (its tile of the smoothing and
                  restriction sweeps), FP64 cycle
  batch_step_cells_fp32  the same for the FP32 cycle

The tuner times candidate settings on synthetic two-phase media and picks the
fastest: first grid_tail_side x grid_graph over a set of grid-path sizes
(total time), then grid_min_n as the largest size at which the persistent
kernel still beats the grid path, then the two step sizes on a batch of
persistent-kernel solves, each in its own precision.  Results go to a JSON table per device
architecture (darcy/tuning/<arch>.json), which capi.Context and the
process-wide context of darcy._core apply when they are created.

    python -m darcy.autotune [--sizes 128 256 512 1024 2048] [--repeats 3]
                             [--batch-n 256] [--batch-count 1000]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Iterable

import numpy as np

from . import capi
from .fields import gaussian_random_field, two_phase_permeability

TABLE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
KEYS = ("grid_min_n", "grid_tail_side", "grid_graph", "batch_step_cells", "batch_step_cells_fp32")
STEP_CELLS = (512, 1024, 2048, 4096, 8192)


def table_path(arch: str) -> str:
    return os.path.join(TABLE_DIR, f"{arch.split(':')[0]}.json")


def load(arch: str) -> dict | None:
    path = table_path(arch)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def apply(ctx: capi.Context, table: dict | None = None) -> dict:
    """Set the tuned values of `table` (default: this device's table) on ctx."""
    if table is None:
        table = load(ctx.info().split(",")[0])
    chosen = (table or {}).get("chosen", {})
    for k in KEYS:
        if k in chosen:
            ctx.set_tuning(k, int(chosen[k]))
    return chosen


def _time_solve(ctx: capi.Context, n: int, a_ptr: int, u_ptr: int, precision: str, repeats: int) -> float:
    ctx.solve_device(n, a_ptr, u_ptr, precision=precision)  # warm-up (and graph capture)
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.solve_device(n, a_ptr, u_ptr, precision=precision)
        best = min(best, time.perf_counter() - t0)
    return best * 1e3


def _time_batch(ctx: capi.Context, n: int, a_ptrs: list, u_ptrs: list, precision: str, repeats: int) -> float:
    ctx.solve_batch_device([n] * len(a_ptrs), a_ptrs, u_ptrs, precision=precision)  # warm-up
    best = float("inf")
    for _ in range(repeats):
        ctx.solve_batch_device([n] * len(a_ptrs), a_ptrs, u_ptrs, precision=precision)
        best = min(best, ctx.last_kernel_ms())
    return best


def tune(ctx: capi.Context, sizes: Iterable[int], small: Iterable[int], repeats: int = 3,
         precision: str = "mixed", log=print, batch_n: int = 256, batch_count: int = 1000) -> dict:
    sizes, small = list(sizes), list(small)
    media = {n: two_phase_permeability(gaussian_random_field(n, seed=n)) for n in set(sizes) | set(small)}
    bufs = {}
    for n, a in media.items():
        pa, pu = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
        ctx.upload(pa, a)
        bufs[n] = (pa, pu)
    saved = {k: ctx.get_tuning(k) for k in KEYS}
    results: dict = {"grid": [], "switch": [], "batch": []}
    batch_a = np.stack([two_phase_permeability(gaussian_random_field(batch_n, seed=i)) for i in range(batch_count)])
    batch_base, batch_u = ctx.alloc(batch_a.nbytes), ctx.alloc(batch_a.nbytes)
    ctx.upload(batch_base, batch_a)
    cells = batch_n * batch_n
    batch_ptrs = [batch_base + 8 * cells * i for i in range(batch_count)]
    batch_out = [batch_u + 8 * cells * i for i in range(batch_count)]
    try:
        # 1. Grid-path shape: every size forced onto the grid path.
        ctx.set_tuning("grid_min_n", 1)
        best = None
        for tail in (32, 64, 128):
            for graph in (0, 1):
                ctx.set_tuning("grid_tail_side", tail)
                ctx.set_tuning("grid_graph", graph)
                ms = {n: _time_solve(ctx, n, *bufs[n], precision, repeats) for n in sizes}
                total = sum(ms.values())
                results["grid"].append({"grid_tail_side": tail, "grid_graph": graph, "ms": ms, "total_ms": total})
                log(json.dumps(results["grid"][-1]))
                if best is None or total < best[0]:
                    best = (total, tail, graph)
        _, tail, graph = best
        ctx.set_tuning("grid_tail_side", tail)
        ctx.set_tuning("grid_graph", graph)
        # 2. Where single solves switch from the persistent kernel to the grid
        #    path: the largest tried size at which the persistent kernel is
        #    still as fast, and never below the smallest tried size minus one
        #    (sizes below it stay on the persistent kernel untimed).
        switch = min(small) - 1
        for n in sorted(small):
            ctx.set_tuning("grid_min_n", n)  # n itself: persistent kernel
            t_batch = _time_solve(ctx, n, *bufs[n], precision, repeats)
            ctx.set_tuning("grid_min_n", n - 1)  # grid path
            t_grid = _time_solve(ctx, n, *bufs[n], precision, repeats)
            results["switch"].append({"n": n, "persistent_ms": t_batch, "grid_ms": t_grid})
            log(json.dumps(results["switch"][-1]))
            if t_batch <= t_grid:
                switch = n
        # 3. The persistent kernel's pipeline step, on a batch of its solves,
        #    for each cycle precision.
        chosen = {"grid_min_n": switch, "grid_tail_side": tail, "grid_graph": graph}
        for key, prec in (("batch_step_cells", "fp64"), ("batch_step_cells_fp32", "mixed")):
            step = None
            for cells_per_step in STEP_CELLS:
                ctx.set_tuning(key, cells_per_step)
                ms = _time_batch(ctx, batch_n, batch_ptrs, batch_out, prec, repeats)
                results["batch"].append({key: cells_per_step, "precision": prec, "n": batch_n,
                                         "count": batch_count, "kernel_ms": ms})
                log(json.dumps(results["batch"][-1]))
                if step is None or ms < step[0]:
                    step = (ms, cells_per_step)
            ctx.set_tuning(key, step[1])
            chosen[key] = step[1]
    finally:
        for k, v in saved.items():
            ctx.set_tuning(k, v)
        for pa, pu in bufs.values():
            ctx.free(pa)
            ctx.free(pu)
        ctx.free(batch_base)
        ctx.free(batch_u)
    return {"device": ctx.info(), "precision": precision, "repeats": repeats, "chosen": chosen, "results": results}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m darcy.autotune")
    ap.add_argument("--sizes", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--small", type=int, nargs="+", default=[65, 96, 128, 192, 256])
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-count", type=int, default=1000)
    ap.add_argument("--out", default=None, help="table path (default: darcy/tuning/<arch>.json)")
    args = ap.parse_args(argv)
    with capi.Context(apply_tuning=False) as ctx:
        table = tune(ctx, args.sizes, args.small, args.repeats, log=lambda s: print(s, flush=True),
                     batch_n=args.batch_n, batch_count=args.batch_count)
        out = args.out or table_path(ctx.info().split(",")[0])
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as fh:
        json.dump(table, fh, indent=1)
    print(json.dumps({"chosen": table["chosen"], "saved": out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
