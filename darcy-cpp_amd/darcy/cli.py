"""``python -m darcy`` -- the command line of the GPU solve path.

Four subcommands, with the reference command's names, flags, defaults,
report lines, ``.npz`` keys and exit codes (python/darcy/cli.py:16-205 and
benchmark.py:54-140 of the reference), so scripts written against it keep
working:

  solve    one medium (generated, or loaded from .npy/.npz) -> report, .npz
  dataset  (a, u) training pairs; the media are generated on the GPU
           (darcy_grf_fields) from the reference's normal draws and solved
           in one persistent launch
  bench    the README table: wall time per grid and preconditioner against
           a sparse direct solve
  info     build and device description (JSON)

Exit codes: 0 success, 1 a solve did not converge, 2 invalid input, 130
interrupted.  ``--precision mixed`` (FP32 cycle inside FP64 flexible CG) is
this package's addition; the batched 256^2 benchmark with the learned warm
start is ``bench.py`` at the repository root.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from collections.abc import Callable, Sequence
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

import darcy


# ----------------------------------------------------------------- options

@dataclass(frozen=True)
class Flag:
    """One command-line option: argparse keywords kept as data."""

    name: str
    kind: Any = None
    default: Any = None
    text: str | None = None
    choices: tuple | None = None
    action: str | None = None
    nargs: str | None = None

    def add_to(self, parser: argparse.ArgumentParser) -> None:
        kw: dict[str, Any] = {}
        for key, value in (("type", self.kind), ("default", self.default), ("help", self.text),
                           ("choices", self.choices), ("action", self.action), ("nargs", self.nargs)):
            if value is not None:
                kw[key] = list(value) if key == "choices" else value
        parser.add_argument(self.name, **kw)


# What reaches darcy.solve / solve_batch.  The reference's five solver flags,
# then --precision.
SOLVER = (
    Flag("--tol", float, 1e-10, "stop at this relative residual ||r|| / ||b||"),
    Flag("--max-iter", int, 10_000, "largest number of CG iterations"),
    Flag("--preconditioner", str, "multigrid", "multigrid, jacobi or none", ("multigrid", "jacobi", "none")),
    Flag("--pre-sweeps", int, 1, "red-black sweeps before each restriction"),
    Flag("--post-sweeps", int, 1, "red-black sweeps after each prolongation"),
    Flag("--cycle-index", int, 2, "coarse visits per level: 1 V-cycle, 2 W-cycle", (1, 2)),
    Flag("--precision", str, "fp64", "fp64, or mixed: FP32 cycle inside FP64 flexible CG", ("fp64", "mixed")),
)

# The random medium: a thresholded Gaussian random field.
MEDIUM = (
    Flag("--alpha", float, 2.0, "spectral decay of the random field (larger: smoother)"),
    Flag("--tau", float, 3.0, "inverse correlation length of the random field"),
    Flag("--f", float, 1.0, "the constant forcing"),
)


def _solver_kwargs(ns: argparse.Namespace, names: Sequence[str]) -> dict[str, Any]:
    return {name: getattr(ns, name) for name in names}


# ----------------------------------------------------------------- reports

def _report(rows: Sequence[tuple[str, Any]]) -> None:
    """Label / value lines, labels left-aligned in 13 columns."""
    for label, value in rows:
        print(f"{label:<13} {value}")


def _medium(n: int, seed: Any, alpha: float, tau: float) -> np.ndarray:
    grf = darcy.gaussian_random_field(n, alpha=alpha, tau=tau, seed=seed)
    return darcy.two_phase_permeability(grf)


# ---------------------------------------------------------------- commands

COMMANDS: dict[str, tuple[str, tuple[Flag, ...], Callable[[argparse.Namespace], int]]] = {}


def command(name: str, summary: str, *flags: Flag):
    def register(fn: Callable[[argparse.Namespace], int]):
        COMMANDS[name] = (summary, flags, fn)
        return fn

    return register


def _load_field(path: Path, key: str) -> np.ndarray | str:
    """The field stored in a .npy, or under `key` in a .npz; a message when
    it is not two-dimensional."""
    stored = np.load(path)
    field = stored[key] if isinstance(stored, np.lib.npyio.NpzFile) else stored
    field = np.asarray(field, dtype=np.float64)
    if field.ndim != 2:
        return f"error: expected a 2-D field in {path}, got shape {field.shape}"
    return field


@command(
    "solve", "solve one problem and print a report",
    Flag("--grid", int, 128, "side of the generated field"),
    Flag("--seed", int, 0, "seed of the generated field"),
    *MEDIUM,
    Flag("--input", Path, None, "read the permeability from this .npy / .npz instead"),
    Flag("--key", str, "a", "array name inside an --input .npz"),
    Flag("--output", Path, None, "save a and u to this .npz"),
    *SOLVER,
)
def run_solve(ns: argparse.Namespace) -> int:
    if ns.input is None:
        a = _medium(ns.grid, ns.seed, ns.alpha, ns.tau)
    else:
        a = _load_field(ns.input, ns.key)
        if isinstance(a, str):
            print(a, file=sys.stderr)
            return 2
    t0 = time.perf_counter()
    res = darcy.solve(a, ns.f, **_solver_kwargs(ns, ("tol", "max_iter", "preconditioner", "pre_sweeps",
                                                     "post_sweeps", "cycle_index", "precision")))
    seconds = time.perf_counter() - t0
    n = a.shape[0]
    rows = [("grid", f"{n} x {n}"), ("preconditioner", f"{ns.preconditioner} ({res.levels} levels)"),
            ("iterations", res.iterations), ("residual", f"{res.residual:.3e}"), ("converged", res.converged),
            ("wall time", f"{seconds:.4f} s"), ("max |u|", f"{np.abs(res.u).max():.6e}")]
    if ns.output is not None:
        ns.output.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(ns.output, a=a, u=res.u)
        rows.append(("wrote", ns.output))
    _report(rows)
    return 0 if res.converged else 1


# Samples whose normals are drawn, turned into media and solved together:
# bounds host and device memory for large datasets.
DATASET_CHUNK_CELLS = 1 << 26


@command(
    "dataset", "generate (a, u) pairs for operator learning",
    Flag("--n-samples", int, 1000, "number of media"),
    Flag("--grid", int, 64, "side of every medium"),
    Flag("--seed", int, 0, "seed of the one generator all media are drawn from"),
    *MEDIUM,
    Flag("--threads", int, 0, "host threads (the solves run as one GPU launch; kept for scripts)"),
    Flag("--dtype", str, "float32", "dtype of the stored arrays", ("float32", "float64")),
    Flag("--output", Path, Path("data/darcy_train.npz"), "destination .npz"),
    *SOLVER,
)
def run_dataset(ns: argparse.Namespace) -> int:
    """All media come from one default_rng(seed) stream, sample after sample
    (the reference draws each sample's n x n normals in turn; a single
    (count, n, n) draw is the same sequence).  The GPU turns the normals
    into fields and media and solves them; like the reference, only the
    tolerance, the iteration cap and the preconditioner are passed to the
    solver (plus --precision)."""
    from darcy import capi

    n, count = ns.grid, ns.n_samples
    if n < 1 or count < 1:
        raise ValueError("--grid and --n-samples must be positive")
    rng = np.random.default_rng(ns.seed)
    media = np.empty((count, n, n), dtype=np.float64)
    sol = np.empty_like(media)
    iters = np.empty(count, dtype=np.int64)
    ok = np.empty(count, dtype=bool)
    step = max(1, DATASET_CHUNK_CELLS // (n * n))
    seconds = 0.0
    with capi.Context() as ctx:
        for s0 in range(0, count, step):
            s1 = min(count, s0 + step)
            xi = rng.standard_normal((s1 - s0, n, n))
            t0 = time.perf_counter()
            _, media[s0:s1] = ctx.grf(xi, alpha=ns.alpha, tau=ns.tau, field=False)
            out = darcy.solve_batch(media[s0:s1], ns.f, tol=ns.tol, max_iter=ns.max_iter,
                                    preconditioner=ns.preconditioner, threads=ns.threads, precision=ns.precision)
            seconds += time.perf_counter() - t0
            sol[s0:s1], iters[s0:s1], ok[s0:s1] = out.u, out.iterations, out.converged
    if not ok.all():
        print(f"error: {int((~ok).sum())}/{count} samples did not converge", file=sys.stderr)
        return 1
    ns.output.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(ns.output, a=media.astype(ns.dtype), u=sol.astype(ns.dtype))
    print(f"wrote {ns.output}: {count} samples at {n}x{n}")
    print(f"  {seconds:.2f} s total, {seconds / count * 1e3:.1f} ms/sample, {iters.mean():.1f} mean CG iterations")
    print(f"  threads: {ns.threads or darcy.num_threads()}")
    return 0


def _min_time(fn: Callable[[], Any], repeats: int) -> tuple[float, Any]:
    """Least wall time over `repeats` calls, and the last result."""
    times, result = [], None
    for _ in range(max(1, repeats)):
        t0 = time.perf_counter()
        result = fn()
        times.append(time.perf_counter() - t0)
    return min(times), result


TABLE_HEAD = ("| grid | preconditioner | CG iters | levels | time (ms) | SciPy direct (ms) | speed-up | rel. err |\n"
              "|---|---|---|---|---|---|---|---|")


def _table_line(row: dict[str, Any]) -> str:
    direct, secs, err = row["scipy_seconds"], row["seconds"], row["relative_error"]
    cells = [row["grid"], row["preconditioner"], row["iterations"], row["levels"], f"{secs * 1e3:.1f}",
             "-" if direct is None else f"{direct * 1e3:.1f}", "-" if direct is None else f"{direct / secs:.2f}x",
             "-" if err is None else f"{err:.1e}"]
    return "| " + " | ".join(str(c) for c in cells) + " |"


@command(
    "bench", "time the solver per grid size against a direct solve",
    Flag("--grids", int, [64, 128, 256, 512], "grid sides", nargs="+"),
    Flag("--repeats", int, 3, "timed repeats per row (the minimum is kept)"),
    Flag("--no-scipy", None, None, "leave out the sparse direct baseline", action="store_true"),
    Flag("--preconditioners", str, ["multigrid", "jacobi"], "preconditioners to time",
         ("multigrid", "jacobi", "none"), nargs="+"),
    Flag("--precision", str, "fp64", "fp64 or mixed", ("fp64", "mixed")),
    Flag("--json", Path, None, "also write the rows to this JSON file"),
)
def run_bench(ns: argparse.Namespace) -> int:
    """Per grid: the seed-0 two-phase medium, the minimum wall time of
    darcy.solve per preconditioner (iteration cap 1000 for multigrid, 100000
    otherwise), and SciPy's sparse direct solve of the same operator as the
    baseline and the error reference."""
    from darcy.direct import solve_direct

    info = darcy.build_info()
    print(f"darcy-cpp {info['version']} | {info['compiler']}")
    print(f"device: {darcy.device_info()}")
    print()
    rows: list[dict[str, Any]] = []
    for n in ns.grids:
        a = _medium(n, 0, 2.0, 3.0)
        direct_s, u_ref = (None, None) if ns.no_scipy else _min_time(lambda: solve_direct(a), 1)
        for pc in ns.preconditioners:
            cap = 1_000 if pc == "multigrid" else 100_000
            secs, res = _min_time(lambda: darcy.solve(a, tol=1e-10, preconditioner=pc, max_iter=cap,
                                                      precision=ns.precision), ns.repeats)
            err = None if u_ref is None else float(np.linalg.norm(res.u - u_ref) / np.linalg.norm(u_ref))
            row = {"grid": n, "preconditioner": pc, "seconds": secs, "iterations": res.iterations,
                   "levels": res.levels, "residual": res.residual, "scipy_seconds": direct_s,
                   "relative_error": err}
            rows.append(row)
            speed = "-" if direct_s is None else f"{direct_s / secs:.2f}x"
            print(f"  {n:>5} {pc:>10}  {secs * 1e3:>9.1f} ms  {res.iterations:>6} iters  speedup {speed:>7}  "
                  f"rel.err {'-' if err is None else f'{err:.1e}'}")
    print()
    print("\n".join([TABLE_HEAD] + [_table_line(r) for r in rows]))
    if ns.json is not None:
        ns.json.parent.mkdir(parents=True, exist_ok=True)
        ns.json.write_text(json.dumps(rows, indent=2))
        print(f"wrote {ns.json}")
    return 0


@command("info", "print build and device information")
def run_info(_: argparse.Namespace) -> int:
    info = dict(darcy.build_info())
    try:
        info["device"] = darcy.device_info()
    except RuntimeError as exc:  # no usable GPU: the build details still print
        info["device"] = f"unavailable: {exc}"
    print(json.dumps(info, indent=2))
    return 0


# -------------------------------------------------------------- entry point

def build_parser() -> argparse.ArgumentParser:
    fmt = argparse.ArgumentDefaultsHelpFormatter
    parser = argparse.ArgumentParser(prog="python -m darcy", formatter_class=fmt,
                                     description="2-D Darcy flow: multigrid-preconditioned CG on an AMD MI355X.")
    parser.add_argument("--version", action="version", version=f"darcy-cpp {darcy.__version__}")
    sub = parser.add_subparsers(dest="command", required=True)
    for name, (summary, flags, fn) in COMMANDS.items():
        p = sub.add_parser(name, help=summary, formatter_class=fmt)
        for flag in flags:
            flag.add_to(p)
        p.set_defaults(func=fn)
    return parser


def main(argv: Sequence[str] | None = None) -> int:
    ns = build_parser().parse_args(argv)
    try:
        return int(ns.func(ns))
    except ValueError as exc:  # input rejected by the solver or the loader
        print(f"error: {exc}", file=sys.stderr)
        return 2
    except KeyboardInterrupt:
        print("interrupted", file=sys.stderr)
        return 130


if __name__ == "__main__":
    raise SystemExit(main())
