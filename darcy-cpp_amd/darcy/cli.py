"""Command line: ``python -m darcy solve | dataset | bench | info``.

The commands, flags, defaults, printed lines, output files and exit codes
are those of the reference's ``darcy`` command (python/darcy/cli.py:16-200,
benchmark.py:54-140); the solves run on the GPU through this package.  One
option is added: ``--precision`` (``fp64``, the reference's arithmetic, or
``mixed``, the FP32 cycle inside FP64 flexible CG).  The headline benchmark
(batched 256^2 solves with the learned warm start) is ``bench.py`` at the
repository root.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from collections.abc import Sequence
from pathlib import Path

import numpy as np

import darcy


def _solver_flags(p: argparse.ArgumentParser) -> None:
    """cli.py:18-29."""
    p.add_argument("--tol", type=float, default=1e-10, help="relative residual tolerance")
    p.add_argument("--max-iter", type=int, default=10_000, help="iteration cap")
    p.add_argument("--preconditioner", choices=["multigrid", "jacobi", "none"], default="multigrid",
                   help="preconditioner (default: multigrid)")
    p.add_argument("--pre-sweeps", type=int, default=1, help="pre-smoothing sweeps per level")
    p.add_argument("--post-sweeps", type=int, default=1, help="post-smoothing sweeps per level")
    p.add_argument("--cycle-index", type=int, default=2, choices=[1, 2], help="1 = V-cycle, 2 = W-cycle")
    p.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                   help="fp64, or the FP32 cycle inside FP64 flexible CG")


def _medium(grid: int, seed, alpha: float, tau: float) -> np.ndarray:
    return darcy.two_phase_permeability(darcy.gaussian_random_field(grid, alpha=alpha, tau=tau, seed=seed))


def cmd_solve(args: argparse.Namespace) -> int:
    """One problem (cli.py:36-75): a generated or loaded field, a report, and
    optionally (a, u) in an .npz.  Exit 0 when converged, 1 otherwise."""
    if args.input is not None:
        loaded = np.load(args.input)
        a = loaded[args.key] if isinstance(loaded, np.lib.npyio.NpzFile) else loaded
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2:
            print(f"error: expected a 2-D field in {args.input}, got shape {a.shape}", file=sys.stderr)
            return 2
    else:
        a = _medium(args.grid, args.seed, args.alpha, args.tau)

    t0 = time.perf_counter()
    r = darcy.solve(a, args.f, tol=args.tol, max_iter=args.max_iter, preconditioner=args.preconditioner,
                    pre_sweeps=args.pre_sweeps, post_sweeps=args.post_sweeps, cycle_index=args.cycle_index,
                    precision=args.precision)
    elapsed = time.perf_counter() - t0

    print(f"grid          {a.shape[0]} x {a.shape[0]}")
    print(f"preconditioner {args.preconditioner} ({r.levels} levels)")
    print(f"iterations    {r.iterations}")
    print(f"residual      {r.residual:.3e}")
    print(f"converged     {r.converged}")
    print(f"wall time     {elapsed:.4f} s")
    print(f"max |u|       {np.abs(r.u).max():.6e}")
    if args.output is not None:
        args.output.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(args.output, a=a, u=r.u)
        print(f"wrote         {args.output}")
    return 0 if r.converged else 1


def cmd_dataset(args: argparse.Namespace) -> int:
    """(a, u) pairs for operator learning (cli.py:77-115): n_samples media
    drawn from one default_rng(seed) stream, one batched solve, arrays in
    ``--dtype``.  As in the reference, only the tolerance, the iteration cap
    and the preconditioner reach the solver.  Exit 1 if any sample fails."""
    rng = np.random.default_rng(args.seed)
    a_all = np.empty((args.n_samples, args.grid, args.grid), dtype=np.float64)
    for i in range(args.n_samples):
        a_all[i] = _medium(args.grid, rng, args.alpha, args.tau)

    t0 = time.perf_counter()
    batch = darcy.solve_batch(a_all, args.f, tol=args.tol, max_iter=args.max_iter,
                              preconditioner=args.preconditioner, threads=args.threads, precision=args.precision)
    elapsed = time.perf_counter() - t0
    if not batch.all_converged:
        failed = int((~batch.converged).sum())
        print(f"error: {failed}/{args.n_samples} samples did not converge", file=sys.stderr)
        return 1

    args.output.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(args.output, a=a_all.astype(args.dtype), u=batch.u.astype(args.dtype))
    per = elapsed / args.n_samples
    print(f"wrote {args.output}: {args.n_samples} samples at {args.grid}x{args.grid}")
    print(f"  {elapsed:.2f} s total, {per * 1e3:.1f} ms/sample, {batch.iterations.mean():.1f} mean CG iterations")
    print(f"  threads: {args.threads or darcy.num_threads()}")
    return 0


def _best_time(fn, repeats: int):
    best, result = float("inf"), None
    for _ in range(repeats):
        t0 = time.perf_counter()
        result = fn()
        best = min(best, time.perf_counter() - t0)
    return best, result


def _table(rows: list[dict]) -> str:
    """The Markdown table of benchmark.py:123-140."""
    lines = ["| grid | preconditioner | CG iters | levels | time (ms) | SciPy direct (ms) | speed-up | rel. err |",
             "|" + "|".join(["---"] * 8) + "|"]
    for r in rows:
        d, t, e = r["scipy_seconds"], r["seconds"], r["relative_error"]
        direct = "-" if d is None else f"{d * 1e3:.1f}"
        speed = "-" if d is None else f"{d / t:.2f}x"
        err = "-" if e is None else f"{e:.1e}"
        lines.append(f"| {r['grid']} | {r['preconditioner']} | {r['iterations']} | {r['levels']} | {t * 1e3:.1f} | "
                     f"{direct} | {speed} | {err} |")
    return "\n".join(lines)


def cmd_bench(args: argparse.Namespace) -> int:
    """The reference's benchmark rows (benchmark.py:54-140) on the GPU: per
    grid the seed-0 two-phase medium, the minimum wall time over repeats of
    darcy.solve per preconditioner (iteration cap 1000 for multigrid, 100000
    otherwise), and a SciPy direct solve of the assembled operator as the
    baseline and the error reference.  Prints the rows and the Markdown table
    of the reference's README; --json writes the rows."""
    from darcy.direct import solve_direct

    info = darcy.build_info()
    print(f"darcy-cpp {info['version']} | {info['compiler']}")
    print(f"device: {darcy.device_info()}")
    print()
    rows = []
    for n in args.grids:
        a = _medium(n, 0, 2.0, 3.0)
        direct_s, u_ref = (None, None) if args.no_scipy else _best_time(lambda a=a: solve_direct(a), 1)
        for name in args.preconditioners:
            cap = 1_000 if name == "multigrid" else 100_000
            secs, r = _best_time(lambda a=a, name=name, cap=cap: darcy.solve(
                a, tol=1e-10, preconditioner=name, max_iter=cap, precision=args.precision), args.repeats)
            err = None if u_ref is None else float(np.linalg.norm(r.u - u_ref) / np.linalg.norm(u_ref))
            row = {"grid": n, "preconditioner": name, "seconds": secs, "iterations": r.iterations,
                   "levels": r.levels, "residual": r.residual, "scipy_seconds": direct_s, "relative_error": err}
            rows.append(row)
            speed = f"{direct_s / secs:.2f}x" if direct_s else "-"
            err_s = "-" if err is None else f"{err:.1e}"
            print(f"  {n:>5} {name:>10}  {secs * 1e3:>9.1f} ms  {r.iterations:>6} iters  speedup {speed:>7}  "
                  f"rel.err {err_s}")
    print()
    print(_table(rows))
    if args.json is not None:
        args.json.parent.mkdir(parents=True, exist_ok=True)
        args.json.write_text(json.dumps(rows, indent=2))
        print(f"wrote {args.json}")
    return 0


def cmd_info(_: argparse.Namespace) -> int:
    info = dict(darcy.build_info())
    try:
        info["device"] = darcy.device_info()
    except RuntimeError as exc:  # no usable GPU: report it, the build details still print
        info["device"] = f"unavailable: {exc}"
    print(json.dumps(info, indent=2))
    return 0


def build_parser() -> argparse.ArgumentParser:
    """cli.py:138-193."""
    parser = argparse.ArgumentParser(prog="python -m darcy",
                                     description="Matrix-free multigrid solver for 2-D Darcy flow (MI355X).",
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument("--version", action="version", version=f"darcy-cpp {darcy.__version__}")
    sub = parser.add_subparsers(dest="command", required=True)
    fmt = parser.formatter_class

    solve = sub.add_parser("solve", help="solve one problem", formatter_class=fmt)
    solve.add_argument("--grid", type=int, default=128, help="grid size when generating a field")
    solve.add_argument("--seed", type=int, default=0, help="random seed for the generated field")
    solve.add_argument("--alpha", type=float, default=2.0, help="GRF smoothness")
    solve.add_argument("--tau", type=float, default=3.0, help="GRF inverse correlation length")
    solve.add_argument("--f", type=float, default=1.0, help="constant forcing term")
    solve.add_argument("--input", type=Path, help="load the permeability field from .npy/.npz")
    solve.add_argument("--key", default="a", help="array key when --input is an .npz")
    solve.add_argument("--output", type=Path, help="write (a, u) to this .npz")
    _solver_flags(solve)
    solve.set_defaults(func=cmd_solve)

    dataset = sub.add_parser("dataset", help="generate a training dataset", formatter_class=fmt)
    dataset.add_argument("--n-samples", type=int, default=1000)
    dataset.add_argument("--grid", type=int, default=64)
    dataset.add_argument("--seed", type=int, default=0)
    dataset.add_argument("--alpha", type=float, default=2.0)
    dataset.add_argument("--tau", type=float, default=3.0)
    dataset.add_argument("--f", type=float, default=1.0)
    dataset.add_argument("--threads", type=int, default=0, help="accepted for compatibility; the GPU takes the batch")
    dataset.add_argument("--dtype", default="float32", choices=["float32", "float64"])
    dataset.add_argument("--output", type=Path, default=Path("data/darcy_train.npz"))
    _solver_flags(dataset)
    dataset.set_defaults(func=cmd_dataset)

    bench = sub.add_parser("bench", help="run the benchmark suite", formatter_class=fmt)
    bench.add_argument("--grids", type=int, nargs="+", default=[64, 128, 256, 512])
    bench.add_argument("--repeats", type=int, default=3)
    bench.add_argument("--no-scipy", action="store_true", help="skip the SciPy direct-solver baseline")
    bench.add_argument("--preconditioners", nargs="+", default=["multigrid", "jacobi"],
                       choices=["multigrid", "jacobi", "none"])
    bench.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    bench.add_argument("--json", type=Path, help="also write the results as JSON")
    bench.set_defaults(func=cmd_bench)

    info = sub.add_parser("info", help="print build, threading and device information")
    info.set_defaults(func=cmd_info)
    return parser


def main(argv: Sequence[str] | None = None) -> int:
    """cli.py:196-205: input validation errors exit 2 with the message."""
    args = build_parser().parse_args(argv)
    try:
        return int(args.func(args))
    except ValueError as exc:
        print(f"error: {exc}", file=sys.stderr)
        return 2
    except KeyboardInterrupt:
        print("interrupted", file=sys.stderr)
        return 130


if __name__ == "__main__":
    raise SystemExit(main())
