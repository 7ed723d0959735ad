"""Matrix-free multigrid solver for 2-D Darcy flow on AMD Instinct MI355X.

Solves

    -div( a(x) grad u(x) ) = f(x)    on (0, 1)^2,    u = 0 on the boundary

with cell-centred finite volumes (harmonic face averaging), preconditioned CG
and a red-black Gauss-Seidel / 2x2-aggregation multigrid cycle -- the method
of the reference darcy-cpp, with the same Python API (python/darcy/__init__.py
of the reference), running as hand-written HIP kernels for gfx950.

Additions, all keyword-only with defaults that keep the reference behaviour:
``u0`` (initial guess), ``precision="mixed"`` (FP32 cycle inside an FP64
flexible CG), :func:`solve_batch_uneven` (problems of different sizes in one
GPU launch) and :func:`device_info`.

The compiled extension is required: importing this package fails loudly when
it is missing, and solving fails loudly when no GPU is present.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Literal, Sequence

import numpy as np

from . import _core
from .fields import gaussian_random_field, two_phase_permeability

__all__ = [
    "BatchResult",
    "SolveResult",
    "__version__",
    "apply_operator",
    "apply_preconditioner",
    "build_info",
    "device_info",
    "gaussian_random_field",
    "has_openmp",
    "num_levels",
    "num_threads",
    "set_num_threads",
    "solve",
    "solve_batch",
    "solve_batch_uneven",
    "two_phase_permeability",
]

__version__: str = _core.__version__

Preconditioner = Literal["multigrid", "jacobi", "none"]
Precision = Literal["fp64", "mixed"]

_DEFAULT_TOL = 1e-10


@dataclass(frozen=True)
class SolveResult:
    """Outcome of a single solve (fields of the reference's SolveResult)."""

    u: np.ndarray
    iterations: int
    residual: float
    converged: bool
    levels: int
    history: np.ndarray = field(default_factory=lambda: np.empty(0))

    def __repr__(self) -> str:
        return (
            f"SolveResult(shape={self.u.shape}, iterations={self.iterations}, "
            f"residual={self.residual:.3e}, converged={self.converged}, levels={self.levels})"
        )


@dataclass(frozen=True)
class BatchResult:
    """Outcome of a batched solve over independent permeability samples."""

    u: np.ndarray
    iterations: np.ndarray
    residuals: np.ndarray
    converged: np.ndarray

    @property
    def all_converged(self) -> bool:
        return bool(self.converged.all())

    def __repr__(self) -> str:
        return (
            f"BatchResult(batch={self.u.shape[0]}, shape={self.u.shape[1:]}, "
            f"mean_iterations={self.iterations.mean():.1f}, "
            f"all_converged={self.all_converged})"
        )


_tuned = False


def _apply_tuning() -> None:
    """Once per process: the autotuner's launch shapes for this device, if a
    table was made (python -m darcy.autotune).  Results never depend on them."""
    global _tuned
    if _tuned:
        return
    from . import autotune

    try:
        arch = _core.device_info().split(",")[0]
    except RuntimeError:  # no device: the call itself reports that (after input checks)
        return
    _tuned = True
    table = autotune.load(arch)
    for key, value in (table or {}).get("chosen", {}).items():
        _core.set_tuning(key, int(value))


def solve(
    a: Any,
    f: Any = 1.0,
    *,
    tol: float = _DEFAULT_TOL,
    max_iter: int = 10_000,
    preconditioner: Preconditioner = "multigrid",
    pre_sweeps: int = 1,
    post_sweeps: int = 1,
    cycle_index: int = 2,
    record_history: bool = False,
    u0: Any = None,
    precision: Precision = "fp64",
) -> SolveResult:
    """Solve ``-div(a grad u) = f`` with homogeneous Dirichlet conditions.

    Same parameters and errors as the reference (ValueError for a non-square
    or non-positive ``a`` and out-of-range parameters).  ``u0`` starts CG
    from an initial guess; ``precision="mixed"`` runs the multigrid cycle in
    FP32 inside an FP64 flexible CG (same stopping rule on the FP64 residual).
    """
    _apply_tuning()
    raw = _core.solve(
        a,
        f,
        tol=tol,
        max_iter=max_iter,
        preconditioner=preconditioner,
        pre_sweeps=pre_sweeps,
        post_sweeps=post_sweeps,
        cycle_index=cycle_index,
        record_history=record_history,
        u0=u0,
        precision=precision,
    )
    return SolveResult(
        u=raw.u,
        iterations=raw.iterations,
        residual=raw.residual,
        converged=raw.converged,
        levels=raw.levels,
        history=np.asarray(raw.history, dtype=np.float64),
    )


def solve_batch(
    a: Any,
    f: Any = 1.0,
    *,
    tol: float = _DEFAULT_TOL,
    max_iter: int = 10_000,
    preconditioner: Preconditioner = "multigrid",
    pre_sweeps: int = 1,
    post_sweeps: int = 1,
    cycle_index: int = 2,
    threads: int = 0,
    precision: Precision = "fp64",
) -> BatchResult:
    """Solve a batch ``a[batch, n, n]`` of independent problems in one launch.

    ``threads`` is accepted for compatibility and ignored: the whole batch is
    one persistent GPU launch.
    """
    _apply_tuning()
    u, iters, residuals, converged = _core.solve_batch(
        a,
        f,
        tol=tol,
        max_iter=max_iter,
        preconditioner=preconditioner,
        pre_sweeps=pre_sweeps,
        post_sweeps=post_sweeps,
        cycle_index=cycle_index,
        threads=threads,
        precision=precision,
    )
    return BatchResult(
        u=u,
        iterations=np.asarray(iters, dtype=np.int64),
        residuals=np.asarray(residuals, dtype=np.float64),
        converged=np.asarray(converged, dtype=bool),
    )


def solve_batch_uneven(
    a: Sequence[Any],
    f: float = 1.0,
    *,
    tol: float = _DEFAULT_TOL,
    max_iter: int = 10_000,
    preconditioner: Preconditioner = "multigrid",
    pre_sweeps: int = 1,
    post_sweeps: int = 1,
    cycle_index: int = 2,
    precision: Precision = "fp64",
) -> tuple[list[np.ndarray], BatchResult]:
    """Solve square problems of different sizes in one persistent launch.

    Returns the list of solutions and a BatchResult whose ``u`` is an object
    array of the same solutions.
    """
    _apply_tuning()
    us, iters, residuals, converged, _levels = _core.solve_batch_uneven(
        list(a),
        float(f),
        tol=tol,
        max_iter=max_iter,
        preconditioner=preconditioner,
        pre_sweeps=pre_sweeps,
        post_sweeps=post_sweeps,
        cycle_index=cycle_index,
        precision=precision,
    )
    holder = np.empty(len(us), dtype=object)
    for i, u in enumerate(us):
        holder[i] = u
    return us, BatchResult(
        u=holder,
        iterations=np.asarray(iters, dtype=np.int64),
        residuals=np.asarray(residuals, dtype=np.float64),
        converged=np.asarray(converged, dtype=bool),
    )


def apply_operator(a: Any, x: Any) -> np.ndarray:
    """Apply the finite-volume operator ``A`` to ``x``, matrix-free."""
    return _core.apply_operator(a, x)


def apply_preconditioner(
    a: Any,
    r: Any,
    *,
    pre_sweeps: int = 1,
    post_sweeps: int = 1,
    cycle_index: int = 2,
    precision: Precision = "fp64",
) -> np.ndarray:
    """Apply one multigrid cycle to ``r``, starting from a zero initial guess."""
    return _core.apply_preconditioner(
        a, r, pre_sweeps=pre_sweeps, post_sweeps=post_sweeps, cycle_index=cycle_index, precision=precision
    )


def num_levels(n: int) -> int:
    """Multigrid depth for an ``n x n`` grid."""
    return _core.num_levels(n)


def num_threads() -> int:
    """Host threads available to the library's host-side loops."""
    return _core.num_threads()


def set_num_threads(threads: int) -> None:
    """Set the host thread count."""
    _core.set_num_threads(threads)


def has_openmp() -> bool:
    """Whether the host layer was compiled with OpenMP support."""
    return _core.has_openmp()


def device_info() -> str:
    """The GPU the solver runs on (architecture, CUs, HIP runtime)."""
    return _core.device_info()


def build_info() -> dict[str, Any]:
    """Compiler, threading, and version details (the reference's keys)."""
    return {
        "version": __version__,
        "compiler": _core.__compiler__,
        "openmp": has_openmp(),
        "max_threads": num_threads(),
        "numpy": np.__version__,
    }
