"""The CPU oracle, pinned before it is trusted as the GPU path's checker.

Each oracle function is checked against something independent of its own
formula:
  * the reference itself (its headers compiled into oracle/_ref by
    oracle/Makefile): bit-exact solutions, cycles, operators, hierarchies;
  * the reference's published iteration counts (README.md:72-76,
    docs/scaling.png) and the committed table of reference iteration counts
    (tests/golden/ref_iterations.json);
  * SciPy: the assembled CSR operator of the reference's test oracle
    (python/darcy/reference.py:29-73, restated here) and a sparse direct
    solve; a DST-I closed form for constant permeability;
  * algebra: dense P^T A P for the Galerkin hierarchy, symmetry and positive
    definiteness of A and M, the error-reduction bound ||I - M A|| < 0.5.
A mistake planted in the oracle's cycle (a swapped post-smoothing colour
order, done here by comparing against a non-symmetric variant) is caught by
the symmetry test, as the reference's test_multigrid.py:94-119 intends.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2, two_phase

import pyoracle

scipy = pytest.importorskip("scipy")
from scipy import sparse  # noqa: E402
from scipy.sparse.linalg import spsolve  # noqa: E402

needs_ref = pytest.mark.skipif(not pyoracle.reference_available(), reason="oracle/_ref not built here")


def harmonic(x, y):
    return 2.0 * x * y / (x + y)


def assemble(a: np.ndarray):
    """CSR matrix of the finite-volume operator (restates reference.py:29-73)."""
    n = a.shape[0]
    ae, aw, an, as_ = (np.zeros((n, n)) for _ in range(4))
    ae[:, :-1] = harmonic(a[:, :-1], a[:, 1:])
    ae[:, -1] = a[:, -1]
    aw[:, 1:] = harmonic(a[:, 1:], a[:, :-1])
    aw[:, 0] = a[:, 0]
    an[:-1, :] = harmonic(a[:-1, :], a[1:, :])
    an[-1, :] = a[-1, :]
    as_[1:, :] = harmonic(a[1:, :], a[:-1, :])
    as_[0, :] = a[0, :]
    idx = np.arange(n * n).reshape(n, n)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [(ae + aw + an + as_).ravel()]
    for r, c, v in ((idx[:, :-1], idx[:, 1:], ae[:, :-1]), (idx[:, 1:], idx[:, :-1], aw[:, 1:]),
                    (idx[:-1, :], idx[1:, :], an[:-1, :]), (idx[1:, :], idx[:-1, :], as_[1:, :])):
        rows.append(r.ravel())
        cols.append(c.ravel())
        vals.append(-v.ravel())
    return sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(n * n, n * n))


def direct(a, f=1.0):
    n = a.shape[0]
    h2 = (1.0 / (n + 1)) ** 2
    rhs = np.full(n * n, f) if np.isscalar(f) else np.asarray(f, float).ravel()
    return spsolve(assemble(a).tocsc(), rhs * h2).reshape(n, n)


# --------------------------------------------------- against the reference

@needs_ref
@pytest.mark.parametrize("n", [1, 2, 5, 17, 31, 33, 64, 100, 128])
@pytest.mark.parametrize("cycle_index", [1, 2])
def test_solutions_bit_exact_with_reference(n, cycle_index, oracle):
    ref = pyoracle.Reference()
    batch = np.stack([two_phase(n, seed=s) for s in range(2)])
    batch[1] = two_phase(n, seed=7, hi=1e3, lo=1e-3)
    # one OpenMP thread: every reduction runs serially on both sides, each
    # library set through the runtime it binds to (another OpenMP runtime
    # loaded into the process, e.g. by torch, must not decide the order)
    threads, ref_threads = oracle.max_threads(), ref.max_threads()
    oracle.set_num_threads(1)
    ref.set_num_threads(1)
    try:
        uo, io, ro, co = oracle.solve_batch(batch, threads=1, cycle_index=cycle_index)
        ur, ir, rr, cr = ref.solve_batch(batch, threads=1, cycle_index=cycle_index)
    finally:
        oracle.set_num_threads(threads)
        ref.set_num_threads(ref_threads)
    assert np.array_equal(io, ir) and np.array_equal(co, cr)
    assert np.array_equal(uo, ur)
    assert np.array_equal(ro, rr)


@needs_ref
@pytest.mark.parametrize("n", [3, 16, 33, 64])
def test_operator_cycle_and_hierarchy_bit_exact_with_reference(n, oracle, rng):
    """No reductions are involved: bit-exact at any thread count."""
    ref = pyoracle.Reference()
    a = two_phase(n, seed=n)
    x = rng.standard_normal((n, n))
    assert np.array_equal(oracle.apply_operator(a, x), ref.apply_operator(a, x))
    for cyc in (1, 2):
        for sw in (1, 2):
            assert np.array_equal(
                oracle.apply_preconditioner(a, x, cycle_index=cyc, pre_sweeps=sw, post_sweeps=sw),
                ref.apply_preconditioner(a, x, cycle_index=cyc, pre_sweeps=sw, post_sweeps=sw))
    for lv, (d, inv) in zip(oracle.hierarchy(a), ref.hierarchy_diag(a)):
        assert np.array_equal(lv["diag"], d) and np.array_equal(lv["inv_diag"], inv)
    assert oracle.num_levels(n) == ref.num_levels(n)


@needs_ref
@pytest.mark.parametrize("min_coarse_n,max_levels", [(4, 32), (16, 32), (1, 3), (1, 1), (6, 2), (100, 32), (0, 0)])
@pytest.mark.parametrize("n", [13, 37, 64])
def test_hierarchy_options_bit_exact_with_reference(n, min_coarse_n, max_levels, oracle, rng):
    """MultigridOptions.min_coarse_n / max_levels (multigrid.hpp:43-46,
    :55-63) and the 8 + 8-sweep coarsest solve (:129-146)."""
    ref = pyoracle.Reference()
    a = two_phase(n, seed=n)
    x = rng.standard_normal((n, n))
    kw = dict(min_coarse_n=min_coarse_n, max_levels=max_levels)
    for cyc in (1, 2):
        assert np.array_equal(oracle.apply_preconditioner(a, x, cycle_index=cyc, **kw),
                              ref.apply_preconditioner(a, x, cycle_index=cyc, **kw))
    threads, ref_threads = oracle.max_threads(), ref.max_threads()
    oracle.set_num_threads(1)
    ref.set_num_threads(1)
    try:
        got = oracle.solve(a, max_iter=3000, **kw)
        want = ref.solve(a, max_iter=3000, **kw)
    finally:
        oracle.set_num_threads(threads)
        ref.set_num_threads(ref_threads)
    assert (got.iterations, got.levels, got.converged) == (want.iterations, want.levels, want.converged)
    assert np.array_equal(got.u, want.u)


@pytest.mark.parametrize("n,min_coarse_n,max_levels,levels", [
    (64, 1, 32, 7), (64, 4, 32, 5), (64, 5, 32, 5), (64, 8, 32, 4), (64, 64, 32, 1), (64, 1, 3, 3),
    (37, 10, 32, 3), (1, 1, 32, 1), (64, 1, 1, 1), (64, 1, 0, 1)])
def test_hierarchy_options_level_counts(n, min_coarse_n, max_levels, levels, oracle):
    """Closed form of the coarsening loop: sides 64 -> 32 -> 16 -> 8 -> 4 ->
    2 -> 1 (37 -> 19 -> 10 -> 5), stopping at the first side <= min_coarse_n
    or at max_levels levels."""
    r = oracle.solve(two_phase(n, seed=1), max_iter=5000, min_coarse_n=min_coarse_n, max_levels=max_levels)
    assert r.levels == levels and r.converged


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_coarsest_sweeps_as_block_gauss_seidel(n, oracle, rng):
    """With one level the cycle is the coarsest solve alone
    (multigrid.hpp:129-146): from x = 0, 8 sweeps red then black, then 8
    black then red.  Restated on the dense matrix A (columns from unit
    vectors): a colour sweep is x_C = D_C^-1 (b_C - A_{C,other} x_other),
    since cells of one colour do not touch each other."""
    a = two_phase(8, seed=3)[:n, :n].copy()
    r = rng.standard_normal((n, n))
    z = oracle.apply_preconditioner(a, r, max_levels=1)
    A = np.stack([oracle.apply_operator(a, e.reshape(n, n)).ravel() for e in np.eye(n * n)], axis=1)
    i, j = np.divmod(np.arange(n * n), n)
    red = (i + j) % 2 == 0
    b, x = r.ravel(), np.zeros(n * n)

    def sweep(c):
        m = red if c == 0 else ~red
        x[m] = (b[m] - A[np.ix_(m, ~m)] @ x[~m]) / np.diag(A)[m]

    for _ in range(8):
        sweep(0)
        sweep(1)
    for _ in range(8):
        sweep(1)
        sweep(0)
    np.testing.assert_allclose(z.ravel(), x, rtol=1e-12, atol=1e-14 * np.abs(x).max())
    # and it approaches the exact solve A^-1 r
    exact = np.linalg.solve(A, b)
    assert np.linalg.norm(x - exact) <= 1e-2 * np.linalg.norm(exact)


@pytest.mark.parametrize("min_coarse_n,max_levels", [(4, 32), (1, 2), (16, 32)])
def test_hierarchy_options_preconditioner_symmetric(min_coarse_n, max_levels, oracle, rng):
    """The cycle stays a symmetric operator with a swept coarsest level."""
    n = 20
    a = two_phase(n, seed=5)
    x, y = rng.standard_normal((2, n, n))
    mx = oracle.apply_preconditioner(a, x, min_coarse_n=min_coarse_n, max_levels=max_levels)
    my = oracle.apply_preconditioner(a, y, min_coarse_n=min_coarse_n, max_levels=max_levels)
    assert abs(np.vdot(y, mx) - np.vdot(x, my)) <= 1e-10 * abs(np.vdot(x, mx))


# -------------------------------------------------- published known answers

@pytest.mark.parametrize("n,iters,levels", [(32, 18, 6), (64, 21, 7), (128, 24, 8), (256, 26, 9), (512, 28, 10),
                                            (1024, 31, 11)])
def test_published_iteration_counts(n, iters, levels, oracle):
    """README.md:72-76 (64..1024) and docs/scaling.png (32): seed 0, W-cycle, 1e-10."""
    r = oracle.solve(two_phase(n, seed=0))
    assert (r.iterations, r.levels, r.converged) == (iters, levels, True)


def test_iteration_table_of_the_reference(oracle):
    """Reference iteration counts (one OpenMP thread) at sizes and media the
    README does not publish, generated by tests/golden/make_golden.py."""
    rows = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_iterations.json")))
    threads = oracle.max_threads()
    oracle.set_num_threads(1)
    try:
        for row in rows:
            hi, lo = (12.0, 3.0) if row["medium"] == "two_phase" else (1e3, 1e-3)
            r = oracle.solve(two_phase(row["n"], seed=row["seed"], hi=hi, lo=lo), cycle_index=row["cycle_index"],
                             max_iter=5000)
            got = (r.iterations, r.levels, r.converged)
            assert got == (row["iterations"], row["levels"], row["converged"]), row
    finally:
        oracle.set_num_threads(threads)


def test_level_count_closed_form(oracle):
    for n in list(range(1, 70)) + [96, 192, 255, 256, 257, 384, 512, 1024, 4096]:
        expected = 1 if n == 1 else int(np.ceil(np.log2(n))) + 1
        assert oracle.num_levels(n) == expected


# -------------------------------------------------------- SciPy oracle

@pytest.mark.parametrize("n", [16, 31, 32, 48, 64])
@pytest.mark.parametrize("preconditioner", ["multigrid", "jacobi", "none"])
def test_matches_sparse_direct_solve(n, preconditioner, oracle):
    a = two_phase(n, seed=n)
    r = oracle.solve(a, tol=1e-12, max_iter=200_000, preconditioner=preconditioner)
    assert r.converged
    assert rel_l2(r.u, direct(a)) < 1e-8


def test_operator_matches_assembled_matrix(oracle, rng):
    n = 17
    a = two_phase(n, seed=11)
    m = assemble(a)
    for _ in range(5):
        x = rng.standard_normal((n, n))
        np.testing.assert_allclose(oracle.apply_operator(a, x), (m @ x.ravel()).reshape(n, n), atol=1e-12)


def test_constant_coefficient_closed_form(oracle):
    """a = c: A = c * (5-point Dirichlet Laplacian), diagonalised by DST-I."""
    from scipy.fft import dstn, idstn

    n, c = 40, 2.5
    h2 = (1.0 / (n + 1)) ** 2
    k = np.arange(1, n + 1)
    lam1 = 2.0 - 2.0 * np.cos(np.pi * k / (n + 1))
    lam = c * (lam1[:, None] + lam1[None, :])
    u_exact = idstn(dstn(np.full((n, n), h2), type=1) / lam, type=1)
    r = oracle.solve(np.full((n, n), c), tol=1e-13)
    assert rel_l2(r.u, u_exact) < 1e-11


def test_variable_forcing(oracle):
    n = 32
    a = two_phase(n, seed=3)
    xs = np.linspace(0, 1, n)
    f = np.outer(np.sin(np.pi * xs), 1.0 + xs**2)
    assert rel_l2(oracle.solve(a, f, tol=1e-12).u, direct(a, f)) < 1e-8


def test_high_contrast(oracle):
    a = two_phase(64, seed=13, hi=1e6, lo=1.0)
    r = oracle.solve(a, tol=1e-10, max_iter=2000)
    assert r.converged and rel_l2(r.u, direct(a)) < 1e-6


# ------------------------------------------------------------ algebra

def _dense(oracle, a):
    n = a.shape[0]
    cols = []
    for k in range(n * n):
        e = np.zeros(n * n)
        e[k] = 1.0
        cols.append(oracle.apply_operator(a, e.reshape(n, n)).ravel())
    return np.array(cols).T


def _dense_from_faces(lv):
    """Dense matrix of a level from the oracle's six face arrays."""
    n = lv["e"].shape[0]
    m = np.zeros((n * n, n * n))
    for i in range(n):
        for j in range(n):
            k = i * n + j
            m[k, k] = lv["diag"][i, j]
            if j + 1 < n:
                m[k, k + 1] = -lv["e"][i, j]
            if j > 0:
                m[k, k - 1] = -lv["w"][i, j]
            if i + 1 < n:
                m[k, k + n] = -lv["n"][i, j]
            if i > 0:
                m[k, k - n] = -lv["s"][i, j]
    return m


@pytest.mark.parametrize("n", [6, 7, 8, 13])
def test_coarsening_is_the_galerkin_product(n, oracle):
    """Reference tests/cpp/test_core.cpp:132-163: coarsen() == P^T A P."""
    a = np.where(np.random.default_rng(8).random((n, n)) < 0.5, 3.0, 12.0)
    levels = oracle.hierarchy(a)
    for fine, coarse in zip(levels[:-1], levels[1:]):
        nf, nc = fine["e"].shape[0], coarse["e"].shape[0]
        p = np.zeros((nf * nf, nc * nc))
        for i in range(nf):
            for j in range(nf):
                p[i * nf + j, (i // 2) * nc + j // 2] = 1.0
        np.testing.assert_allclose(p.T @ _dense_from_faces(fine) @ p, _dense_from_faces(coarse), atol=1e-12)


def test_operator_symmetric_positive_definite(oracle, rng):
    a = two_phase(12, seed=5)
    m = _dense(oracle, a)
    np.testing.assert_allclose(m, m.T, atol=1e-12)
    assert np.linalg.eigvalsh(m).min() > 0


@pytest.mark.parametrize("cycle_index", [1, 2])
@pytest.mark.parametrize("sweeps", [1, 2, 3])
@pytest.mark.parametrize("fp32", [False, True])
def test_preconditioner_symmetric(cycle_index, sweeps, fp32, oracle):
    rng = np.random.default_rng(4)
    a = two_phase(32, seed=8)
    tol = 1e-5 if fp32 else 1e-10
    for _ in range(3):
        r, s = rng.standard_normal((2, 32, 32))
        mr = oracle.apply_preconditioner(a, r, pre_sweeps=sweeps, post_sweeps=sweeps, cycle_index=cycle_index,
                                         fp32=fp32)
        ms = oracle.apply_preconditioner(a, s, pre_sweeps=sweeps, post_sweeps=sweeps, cycle_index=cycle_index,
                                         fp32=fp32)
        left, right = float(np.sum(mr * s)), float(np.sum(ms * r))
        assert abs(left - right) / max(abs(left), abs(right)) < tol


def test_symmetry_test_catches_a_swapped_smoother(oracle):
    """Planted mistake: a cycle whose post-smoothing repeats the pre-smoothing
    colour order (red then black) is not symmetric.  Built from the oracle's
    own pieces: M_bad = S_rb^T-free composition via two half cycles."""
    rng = np.random.default_rng(9)
    n = 16
    a = two_phase(n, seed=2)
    A = _dense(oracle, a)
    D = np.diag(np.diag(A))
    red = ((np.arange(n)[:, None] + np.arange(n)[None, :]) % 2 == 0).ravel()
    L = np.tril(A, -1)

    def gs(colour_first):  # one red/black sweep as an operator x -> x + B (b - A x)
        order = [red, ~red] if colour_first == "red" else [~red, red]
        def sweep(x, b):
            for mask in order:
                xn = x.copy()
                v = b - (A - D) @ x
                xn[mask] = v[mask] / np.diag(A)[mask]
                x = xn
            return x
        return sweep

    def two_grid(b, post_colour):
        x = gs("red")(np.zeros_like(b), b)
        x = gs(post_colour)(x, b)
        return x

    r, s = rng.standard_normal((2, n * n))
    scale = np.linalg.norm(r) * np.linalg.norm(s)
    good = abs(np.dot(two_grid(r, "black"), s) - np.dot(two_grid(s, "black"), r))
    bad = abs(np.dot(two_grid(r, "red"), s) - np.dot(two_grid(s, "red"), r))
    assert good < 1e-13 * scale
    assert bad > 1e-6 * scale
    del L


@pytest.mark.parametrize("cycle_index", [1, 2])
def test_preconditioner_positive_definite_and_effective(cycle_index, oracle):
    rng = np.random.default_rng(5)
    a = two_phase(64, seed=10)
    worst = 0.0
    for _ in range(5):
        x = rng.standard_normal((64, 64))
        assert float(np.sum(oracle.apply_preconditioner(a, x, cycle_index=cycle_index) * x)) > 0.0
        e = x - oracle.apply_preconditioner(a, oracle.apply_operator(a, x), cycle_index=cycle_index)
        worst = max(worst, float(np.linalg.norm(e) / np.linalg.norm(x)))
    if cycle_index == 2:
        assert worst < 0.5


def test_fp32_cycle_is_a_faithful_rounding_of_the_fp64_cycle(oracle, rng):
    a = two_phase(64, seed=12)
    r = rng.standard_normal((64, 64))
    z64 = oracle.apply_preconditioner(a, r)
    z32 = oracle.apply_preconditioner(a, r, fp32=True)
    assert 0 < rel_l2(z32, z64) < 1e-5


def test_flexible_cg_with_fp32_cycle_reaches_the_fp64_solution(oracle):
    a = two_phase(128, seed=1)
    r64 = oracle.solve(a)
    r32 = oracle.solve(a, fp32=True)
    assert r32.converged and abs(r32.iterations - r64.iterations) <= 2
    assert rel_l2(r32.u, r64.u) < 1e-8


def test_warm_start(oracle):
    a = two_phase(64, seed=6)
    ref = oracle.solve(a, tol=1e-13)
    assert oracle.solve(a, u0=ref.u).iterations == 0
    near = oracle.solve(a, u0=ref.u * (1 + 1e-4))
    cold = oracle.solve(a)
    assert near.converged and near.iterations < cold.iterations
    assert rel_l2(near.u, ref.u) < 1e-9


def test_residual_is_honest(oracle):
    n = 64
    a = two_phase(n, seed=2)
    r = oracle.solve(a, tol=1e-11)
    b = np.full((n, n), (1.0 / (n + 1)) ** 2)
    true = np.linalg.norm(b - oracle.apply_operator(a, r.u)) / np.linalg.norm(b)
    assert true < 1e-10 and r.residual == pytest.approx(true, rel=0.05, abs=1e-13)


def test_zero_forcing_and_max_iter(oracle):
    r = oracle.solve(two_phase(16, seed=1), f=0.0)
    assert r.converged and r.iterations == 0 and np.all(r.u == 0)
    r = oracle.solve(two_phase(128, seed=0), tol=1e-14, max_iter=2)
    assert not r.converged and r.iterations == 2


def test_invalid_permeability(oracle):
    a = np.full((8, 8), 2.0)
    a[3, 4] = np.nan
    with pytest.raises(ValueError, match="index=28"):
        oracle.solve(a)
