"""Assembled-matrix baseline for the benchmark command: the 5-point finite
volume operator as a SciPy CSR matrix, solved by SuperLU.

The operator is the solver's (harmonic-mean faces 2xy/(x+y) between cells, a
boundary face taking the cell's own permeability, ghost values zero;
level.hpp:57, 79-119); b = f h^2 with h = 1/(n+1) (solver.hpp:101-102).  It
serves as the direct-solver column of the benchmark table, as the
reference's assembled reference does (python/darcy/reference.py:29-85).
"""

from __future__ import annotations

import numpy as np


def _harmonic(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    return 2.0 * x * y / (x + y)


def assemble(a: np.ndarray):
    """The n^2 x n^2 CSR matrix of the operator for permeability a (n x n)."""
    from scipy import sparse

    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    e = a.copy()
    w = a.copy()
    no = a.copy()
    so = a.copy()
    e[:, :-1] = _harmonic(a[:, :-1], a[:, 1:])
    w[:, 1:] = _harmonic(a[:, 1:], a[:, :-1])
    no[:-1, :] = _harmonic(a[:-1, :], a[1:, :])
    so[1:, :] = _harmonic(a[1:, :], a[:-1, :])
    k = np.arange(n * n).reshape(n, n)
    rows = [k.ravel()]
    cols = [k.ravel()]
    vals = [(e + w + no + so).ravel()]
    for coef, sl_from, sl_to in (
        (e, (slice(None), slice(None, -1)), (slice(None), slice(1, None))),
        (w, (slice(None), slice(1, None)), (slice(None), slice(None, -1))),
        (no, (slice(None, -1), slice(None)), (slice(1, None), slice(None))),
        (so, (slice(1, None), slice(None)), (slice(None, -1), slice(None))),
    ):
        rows.append(k[sl_from].ravel())
        cols.append(k[sl_to].ravel())
        vals.append(-coef[sl_from].ravel())
    return sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(n * n, n * n))


def solve_direct(a: np.ndarray, f: float = 1.0) -> np.ndarray:
    """u for -div(a grad u) = f by a sparse direct solve."""
    from scipy.sparse.linalg import spsolve

    n = np.asarray(a).shape[0]
    b = np.full(n * n, f / float(n + 1) ** 2)
    return spsolve(assemble(a).tocsc(), b).reshape(n, n)
