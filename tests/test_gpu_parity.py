"""GPU solve path against the CPU oracle.

The oracle (oracle/darcy_oracle.cpp) restates the reference arithmetic and is
itself pinned bit-for-bit against the reference headers and the reference's
published iteration counts (tests/test_oracle.py).  Here the HIP kernels are
compared with it on the same inputs:

* apply_operator and the FP64 cycle (apply_preconditioner) are bit-exact:
  the per-cell operations run in the reference's order with FMA contraction
  off, and the cycle contains no reductions.
* The FP32 cycle is bit-exact against the oracle's FP32 cycle.
* Full solves differ only through the summation order of the CG dot
  products: identical iteration counts, u within 1e-9 relative L2.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2, two_phase

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 33, 64])
def test_apply_operator_bit_exact(n, oracle, rng):
    a = two_phase(n, seed=n)
    x = rng.standard_normal((n, n))
    assert np.array_equal(darcy.apply_operator(a, x), oracle.apply_operator(a, x))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 31, 32, 48, 64, 100])
@pytest.mark.parametrize("cycle_index", [1, 2])
def test_fp64_cycle_bit_exact(n, cycle_index, oracle, rng):
    a = two_phase(n, seed=n + 1)
    r = rng.standard_normal((n, n))
    got = darcy.apply_preconditioner(a, r, cycle_index=cycle_index)
    want = oracle.apply_preconditioner(a, r, cycle_index=cycle_index)
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max():.3e}"


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_fp64_cycle_bit_exact_sweeps(sweeps, oracle, rng):
    n = 37
    a = two_phase(n, seed=3)
    r = rng.standard_normal((n, n))
    got = darcy.apply_preconditioner(a, r, pre_sweeps=sweeps, post_sweeps=sweeps)
    want = oracle.apply_preconditioner(a, r, pre_sweeps=sweeps, post_sweeps=sweeps)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [1, 4, 17, 64, 129])
@pytest.mark.parametrize("cycle_index", [1, 2])
def test_fp32_cycle_bit_exact(n, cycle_index, oracle, rng):
    a = two_phase(n, seed=2 * n)
    r = rng.standard_normal((n, n))
    got = darcy.apply_preconditioner(a, r, cycle_index=cycle_index, precision="mixed")
    want = oracle.apply_preconditioner(a, r, cycle_index=cycle_index, fp32=True)
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max():.3e}"


@pytest.mark.parametrize("n,iters,levels", [(32, 18, 6), (64, 21, 7), (128, 24, 8), (256, 26, 9)])
def test_published_iteration_counts(n, iters, levels, oracle):
    """Seed-0 two-phase media, W-cycle, tol 1e-10 (reference README.md:72-76)."""
    a = two_phase(n, seed=0)
    r = darcy.solve(a)
    assert r.converged and r.iterations == iters and r.levels == levels
    ref = oracle.solve(a)
    assert rel_l2(r.u, ref.u) <= 1e-9


@pytest.mark.parametrize("cycle_index", [1, 2])
@pytest.mark.parametrize("preconditioner", ["multigrid", "jacobi", "none"])
def test_solve_matches_oracle(cycle_index, preconditioner, oracle):
    n = 48
    a = two_phase(n, seed=5)
    kw = dict(tol=1e-10, cycle_index=cycle_index, preconditioner=preconditioner, max_iter=20000)
    got = darcy.solve(a, record_history=True, **kw)
    want = oracle.solve(a, record_history=True, **kw)
    assert got.iterations == want.iterations
    assert got.converged == want.converged
    assert rel_l2(got.u, want.u) <= 1e-9
    assert got.history.shape == want.history.shape
    # Preconditioned histories agree closely; plain CG runs hundreds of
    # iterations and its late residuals drift with the dot-product rounding.
    head = 20 if preconditioner == "none" else len(want.history)
    np.testing.assert_allclose(got.history[:head], want.history[:head], rtol=1e-6, atol=1e-15)
    assert got.history[-1] < 1e-10


def test_mixed_precision_solution(oracle):
    n = 128
    a = two_phase(n, seed=11)
    got = darcy.solve(a, precision="mixed")
    ref = oracle.solve(a)
    fp32_oracle = oracle.solve(a, fp32=True)
    assert got.converged and got.residual < 1e-10
    assert abs(got.iterations - fp32_oracle.iterations) <= 1
    assert rel_l2(got.u, ref.u) <= 1e-8


@pytest.mark.parametrize("cap", [1, 2, 3])
def test_mixed_precision_capped_iterations(oracle, rng, cap):
    """The FP32-cycle kernel applies each u += alpha p one pass late; a solve
    stopped by the iteration cap must still carry the last one.  Without it u
    would be off by tens of percent; the two sides differ only in dot-product
    rounding.  Cold and warm starts."""
    a = two_phase(64, seed=5)
    for u0 in (None, 1e-3 * rng.standard_normal((64, 64))):
        got = darcy.solve(a, precision="mixed", tol=1e-14, max_iter=cap, u0=u0)
        want = oracle.solve(a, tol=1e-14, max_iter=cap, fp32=True, u0=u0)
        assert got.iterations == want.iterations == cap and not got.converged
        assert rel_l2(got.u, want.u) <= 1e-6


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_uneven_warm_batch_matches_oracle(oracle, rng, precision):
    """Uneven sizes (tiny, odd, even) from random warm starts in one launch:
    the warm-start init pass, the hierarchy setup at odd sizes and, with the
    FP32 cycle, the deferred u update; each problem as the oracle solves it
    (same cycle arithmetic; dot products summed in another order)."""
    from darcy import capi

    sizes = [1, 2, 3, 5, 17, 33, 64]
    media = [two_phase(n, seed=40 + i) for i, n in enumerate(sizes)]
    u0s = [1e-3 * rng.standard_normal((n, n)) for n in sizes]
    with capi.Context() as ctx:
        bufs = [(ctx.alloc(a.nbytes), ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)) for a in media]
        try:
            for (pa, _, p0), a, u0 in zip(bufs, media, u0s):
                ctx.upload(pa, a)
                ctx.upload(p0, u0)
            res = ctx.solve_batch_device(sizes, [b[0] for b in bufs], [b[1] for b in bufs],
                                         u0_ptrs=[b[2] for b in bufs], precision=precision)
            us = []
            for (_, pu, _), a in zip(bufs, media):
                u = np.empty_like(a)
                ctx.download(u, pu)
                us.append(u)
        finally:
            for b in bufs:
                for p in b:
                    ctx.free(p)
    assert res["converged"].all()
    for a, u0, u, it in zip(media, u0s, us, res["iterations"]):
        want = oracle.solve(a, u0=u0, fp32=precision == "mixed")
        assert abs(int(it) - want.iterations) <= 1
        assert rel_l2(u, want.u) <= 1e-9


def test_uneven_batch_matches_oracle(oracle):
    sizes = [64, 17, 96, 33, 128, 5, 64]
    media = [two_phase(n, seed=i) for i, n in enumerate(sizes)]
    us, res = darcy.solve_batch_uneven(media)
    assert res.all_converged
    for a, u, it in zip(media, us, res.iterations):
        ref = oracle.solve(a)
        assert it == ref.iterations
        assert rel_l2(u, ref.u) <= 1e-9


def test_batch_equals_single_solves():
    batch_a = np.stack([two_phase(32, seed=s) for s in range(6)])
    batch = darcy.solve_batch(batch_a, tol=1e-12)
    assert batch.all_converged
    for i, a in enumerate(batch_a):
        single = darcy.solve(a, tol=1e-12)
        assert single.iterations == batch.iterations[i]
        assert np.array_equal(single.u, batch.u[i])


def test_warm_start_from_solution_converges_immediately(oracle):
    a = two_phase(64, seed=3)
    ref = oracle.solve(a, tol=1e-12)
    r = darcy.solve(a, u0=ref.u, tol=1e-10)
    assert r.converged and r.iterations == 0


def test_warm_start_matches_oracle(oracle, rng):
    a = two_phase(64, seed=4)
    u0 = 1e-3 * rng.standard_normal((64, 64))
    got = darcy.solve(a, u0=u0)
    want = oracle.solve(a, u0=u0)
    assert got.iterations == want.iterations
    assert rel_l2(got.u, want.u) <= 1e-9


def test_invalid_permeability_reports_first_index():
    a = np.full((8, 8), 2.0)
    a[3, 4] = -1.0
    a[5, 1] = 0.0
    with pytest.raises(ValueError, match="positive.*flat index 28"):
        darcy.solve(a)


def test_zero_forcing():
    r = darcy.solve(two_phase(16, seed=1), f=0.0)
    assert r.converged and r.iterations == 0 and np.all(r.u == 0.0)


def test_max_iter_reported_not_raised():
    r = darcy.solve(two_phase(128, seed=0), tol=1e-14, max_iter=2)
    assert not r.converged and r.iterations == 2 and np.isfinite(r.u).all()
