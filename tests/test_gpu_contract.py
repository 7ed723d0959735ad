"""Behaviour of the GPU path at its boundary: the reference's API suite
(tests/test_api.py of the reference), the breakdown exit of solver.hpp:
149-155 forced through the C ABI, bounded history writes, argument checks
made before any launch, and hierarchies with non-default min_coarse_n /
max_levels (multigrid.hpp:43-46, :129-146) against the oracle.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2, two_phase

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")
from darcy import capi  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    with capi.Context() as c:
        yield c


# ------------------------------------------------ reference tests/test_api.py


class TestInputHandling:
    """test_api.py:12-51: awkward but valid inputs are converted."""

    def test_accepts_float32(self):
        a = two_phase(32, seed=0)
        assert np.allclose(darcy.solve(a.astype(np.float32)).u, darcy.solve(a).u, rtol=1e-6)

    def test_accepts_integer_dtype(self):
        assert darcy.solve(np.full((16, 16), 4, dtype=np.int32)).converged

    def test_accepts_fortran_order(self):
        a = two_phase(32, seed=1)
        assert np.array_equal(darcy.solve(np.asfortranarray(a)).u, darcy.solve(a).u)

    def test_accepts_non_contiguous_view(self):
        a = two_phase(64, seed=2)[::2, ::2]
        r = darcy.solve(a)
        assert r.converged and np.array_equal(r.u, darcy.solve(np.ascontiguousarray(a)).u)

    def test_accepts_nested_lists(self):
        assert darcy.solve([[1.0, 2.0], [2.0, 1.0]]).converged

    def test_does_not_mutate_its_input(self):
        a = two_phase(96, seed=3)  # grid path
        b = two_phase(32, seed=3)  # persistent kernel
        ca, cb = a.copy(), b.copy()
        darcy.solve(a)
        darcy.solve(b, np.ones_like(b), u0=np.zeros_like(b))
        darcy.solve_batch(np.stack([b, b]))
        assert np.array_equal(a, ca) and np.array_equal(b, cb)

    def test_output_is_a_fresh_writable_array(self):
        for n in (16, 80):
            result = darcy.solve(two_phase(n, seed=4))
            assert result.u.flags.writeable and result.u.flags.c_contiguous
            result.u[0, 0] = 123.0
            assert darcy.solve(two_phase(n, seed=4)).u[0, 0] != 123.0
        batch = darcy.solve_batch(np.stack([two_phase(16, seed=s) for s in range(2)]))
        assert batch.u.flags.writeable and batch.u.flags.c_contiguous

    def test_scalar_and_uniform_array_forcing_agree(self):
        a = two_phase(24, seed=5)
        u_scalar = darcy.solve(a, 2.5, tol=1e-12).u
        u_array = darcy.solve(a, np.full((24, 24), 2.5), tol=1e-12).u
        assert np.allclose(u_scalar, u_array, atol=1e-13)


class TestBatchBehaviour:
    """test_api.py:110-152."""

    def test_propagates_validation_errors(self):
        batch_a = np.stack([np.full((8, 8), 2.0) for _ in range(3)])
        batch_a[1, 0, 0] = -1.0
        with pytest.raises(ValueError, match="positive"):
            darcy.solve_batch(batch_a)

    def test_uneven_batch_names_the_first_bad_problem(self):
        media = [two_phase(n, seed=n) for n in (16, 24, 40)]
        media[2][1, 1] = np.nan
        media[1][0, 3] = 0.0
        with pytest.raises(ValueError, match=r"flat index 3 of problem 1"):
            darcy.solve_batch_uneven(media)

    def test_single_sample_batch(self):
        a = two_phase(16, seed=1)
        batch = darcy.solve_batch(a[None], tol=1e-12)
        assert np.array_equal(batch.u[0], darcy.solve(a, tol=1e-12).u)

    def test_reports_per_sample_diagnostics(self):
        batch = darcy.solve_batch(np.stack([two_phase(32, seed=s) for s in range(5)]), max_iter=2)
        assert not batch.all_converged
        assert batch.iterations.shape == (5,) and np.all(batch.iterations == 2)
        assert batch.residuals.shape == (5,) and np.all(batch.residuals > 1e-10)
        assert batch.converged.dtype == bool

    def test_threads_argument_does_not_change_results(self):
        batch_a = np.stack([two_phase(48, seed=s) for s in range(4)])
        single = darcy.solve_batch(batch_a, tol=1e-12, threads=1)
        multi = darcy.solve_batch(batch_a, tol=1e-12, threads=8)
        assert np.array_equal(single.iterations, multi.iterations)
        assert np.array_equal(single.u, multi.u)


class TestIntrospection:
    """test_api.py:171-178."""

    def test_result_repr_is_informative(self):
        text = repr(darcy.solve(two_phase(16, seed=0)))
        for token in ("SolveResult", "iterations", "residual", "converged"):
            assert token in text

    def test_history_is_empty_unless_requested(self):
        assert darcy.solve(two_phase(16, seed=0)).history.size == 0
        assert darcy.solve(two_phase(16, seed=0), record_history=True).history.size > 0


# ----------------------------------------------------- the breakdown exit


@pytest.mark.parametrize("n", [40, 130])  # persistent kernel, grid path
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", [1, 4])
def test_forced_breakdown_returns_like_the_reference(ctx, oracle, n, precision, k):
    """solver.hpp:149-155: <p, Ap> of iteration k reads as 0, so the solve
    returns converged = False with iterations = k - 1 and the u and residual
    of the k - 1 iterations before it (the same bits as max_iter = k - 1)."""
    a = two_phase(n, seed=11)
    # The oracle restates the FP32 cycle with its dense tail over the levels
    # of side <= 8; the grid path's default (16) rounds intermediate iterates
    # differently.
    ctx.set_tuning("grid_dense_side", 8)
    try:
        _forced_breakdown(ctx, oracle, a, precision, k)
    finally:
        ctx.set_tuning("grid_dense_side", 16)


def _forced_breakdown(ctx, oracle, a, precision, k):
    got = ctx.solve(a, precision=precision, breakdown_at=k, record_history=True)
    assert got["status"] == 3 and not got["converged"]
    assert got["iterations"] == k - 1
    assert len(got["history"]) == k
    if k == 1:
        assert np.all(got["u"] == 0.0) and got["residual"] == 1.0
        return
    capped = ctx.solve(a, precision=precision, max_iter=k - 1, record_history=True)
    assert capped["status"] == 2 and capped["iterations"] == k - 1
    assert np.array_equal(got["u"], capped["u"]) and got["residual"] == capped["residual"]
    np.testing.assert_array_equal(got["history"], capped["history"])
    want = oracle.solve(a, max_iter=k - 1, fp32=precision == "mixed")
    assert rel_l2(got["u"], want.u) <= 1e-9
    # The breakdown leaves the context usable and the next solve unaffected.
    again = ctx.solve(a, precision=precision)
    assert again["converged"] and again["iterations"] == oracle.solve(a, fp32=precision == "mixed").iterations


def test_breakdown_option_is_checked(ctx):
    with pytest.raises(ValueError, match="breakdown_at"):
        ctx.solve(two_phase(8, seed=0), breakdown_at=-1)


# ------------------------------------------------------- history bounds


@pytest.mark.parametrize("n", [40, 130])
def test_history_cap_bounds_host_writes(ctx, n):
    a = two_phase(n, seed=2)
    full = ctx.solve(a, record_history=True)
    short = ctx.solve(a, record_history=True, history_cap=3)
    assert len(short["history"]) == 3
    np.testing.assert_array_equal(short["history"], full["history"][:3])
    assert short["iterations"] == full["iterations"]


@pytest.mark.parametrize("n", [40, 130])
def test_history_cap_bounds_device_writes(ctx, n):
    """With device pointers the kernels write the caller's buffer directly:
    nothing lands past history_cap (a guard region after it stays intact)."""
    a = two_phase(n, seed=2)
    cap, guard = 4, 64
    sentinel = np.full(cap + guard, -7.0)
    pa, pu, ph = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes), ctx.alloc(sentinel.nbytes)
    try:
        ctx.upload(pa, a)
        ctx.upload(ph, sentinel)
        o = ctx._opts(record_history=True, device_pointers=True)
        st = capi.Stats()
        ctx._check(ctx.lib.darcy_solve(ctx.ptr, C.c_void_p(pa), None, C.c_double(1.0), C.c_int64(n), C.byref(o),
                                       None, C.c_void_p(pu), C.byref(st),
                                       C.cast(C.c_void_p(ph), C.POINTER(C.c_double)), cap))
        back = np.empty_like(sentinel)
        ctx.download(back, ph)
    finally:
        for p in (pa, pu, ph):
            ctx.free(p)
    assert st.converged and st.history_len == cap
    full = ctx.solve(a, record_history=True)["history"]
    np.testing.assert_array_equal(back[:cap], full[:cap])
    assert np.all(back[cap:] == -7.0)


# ------------------------------------------- host pointers shared by problems


def _solve_batch_host(ctx, media, rhs, u0=None, precision="fp64"):
    count = len(media)
    o = ctx._opts(precision=precision)
    sizes = (C.c_int64 * count)(*[m.shape[0] for m in media])
    a = (C.c_void_p * count)(*[m.ctypes.data for m in media])
    r = (C.c_void_p * count)(*[x.ctypes.data if x is not None else None for x in rhs])
    z = (C.c_void_p * count)(*[x.ctypes.data if x is not None else None for x in u0]) if u0 else None
    outs = [np.empty_like(m) for m in media]
    u = (C.c_void_p * count)(*[x.ctypes.data for x in outs])
    st = (capi.Stats * count)()
    ctx._check(ctx.lib.darcy_solve_batch(ctx.ptr, count, sizes, a, r, None, C.byref(o), z, u, st))
    return outs, st


def test_pointer_shared_by_problems_of_different_sizes(ctx, rng):
    """One host rhs / u0 array read as 64^2 by one problem and 128^2 by
    another (first use the smaller) is staged at the larger size."""
    big_f = 1.0 + rng.random((128, 128))
    big_u0 = 1e-4 * rng.standard_normal((128, 128))
    media = [two_phase(64, seed=1), two_phase(128, seed=2)]
    outs, st = _solve_batch_host(ctx, media, [big_f, big_f], [big_u0, big_u0])
    assert all(s.converged for s in st)
    small_f = big_f.reshape(-1)[: 64 * 64].reshape(64, 64)
    small_u0 = big_u0.reshape(-1)[: 64 * 64].reshape(64, 64)
    for m, f, u0, got in zip(media, [small_f, big_f], [small_u0, big_u0], outs):
        # each problem alone, with arrays of its own size (same kernel: same bits)
        alone, _ = _solve_batch_host(ctx, [m], [np.ascontiguousarray(f)], [np.ascontiguousarray(u0)])
        assert np.array_equal(got, alone[0])


# --------------------------------------------- checks before any launch


@pytest.mark.parametrize("device_pointers", [False, True])
def test_sizes_and_null_arrays_rejected_before_launch(ctx, device_pointers):
    tiny = np.ones((2, 2))
    o = ctx._opts(device_pointers=device_pointers)
    st = capi.Stats()
    out = np.empty((2, 2))
    for n, a, u in [(32769, tiny, out), (1 << 20, tiny, out), (128, None, out), (128, tiny, None),
                    (16, None, out), (16, tiny, None), (0, tiny, out)]:
        rc = ctx.lib.darcy_solve(ctx.ptr, a.ctypes.data if a is not None else None, None, C.c_double(1.0),
                                 C.c_int64(n), C.byref(o), None, u.ctypes.data if u is not None else None,
                                 C.byref(st), None, 0)
        assert rc == capi.INVALID_INPUT, (n, a is None, u is None)
        assert ctx.lib.darcy_last_error(ctx.ptr).decode()
    # The context still solves.
    assert ctx.solve(two_phase(16, seed=0))["converged"]


# ----------------------------------- hierarchy options (multigrid.hpp:43-46)

HIERARCHIES = [(4, 32), (16, 32), (1, 3), (1, 1), (6, 2), (100, 32)]


@pytest.mark.parametrize("min_coarse_n,max_levels", HIERARCHIES)
@pytest.mark.parametrize("n", [37, 64])
def test_hierarchy_options_cycle_bit_exact(ctx, oracle, rng, n, min_coarse_n, max_levels):
    """The FP64 cycle with a coarsest level wider than 1 x 1 (8 + 8
    symmetric sweeps, multigrid.hpp:129-146) gives the oracle's bits."""
    a = two_phase(n, seed=n + min_coarse_n)
    r = rng.standard_normal((n, n))
    for cyc in (1, 2):
        got = ctx.apply_preconditioner(a, r, cycle_index=cyc, min_coarse_n=min_coarse_n, max_levels=max_levels)
        want = oracle.apply_preconditioner(a, r, cycle_index=cyc, min_coarse_n=min_coarse_n,
                                           max_levels=max_levels)
        assert np.array_equal(got, want)
        got32 = ctx.apply_preconditioner(a, r, cycle_index=cyc, precision="mixed", min_coarse_n=min_coarse_n,
                                         max_levels=max_levels)
        want32 = oracle.apply_preconditioner(a, r, cycle_index=cyc, fp32=True, min_coarse_n=min_coarse_n,
                                             max_levels=max_levels)
        assert np.array_equal(got32, want32)


@pytest.mark.parametrize("min_coarse_n,max_levels", HIERARCHIES)
@pytest.mark.parametrize("n", [48, 160])  # persistent kernel, grid path
def test_hierarchy_options_solve_matches_oracle(ctx, oracle, n, min_coarse_n, max_levels):
    a = two_phase(n, seed=7)
    got = ctx.solve(a, min_coarse_n=min_coarse_n, max_levels=max_levels, max_iter=4000)
    want = oracle.solve(a, min_coarse_n=min_coarse_n, max_levels=max_levels, max_iter=4000)
    assert got["levels"] == want.levels
    assert got["converged"] and abs(got["iterations"] - want.iterations) <= 1
    assert rel_l2(got["u"], want.u) <= 1e-9
    mixed = ctx.solve(a, min_coarse_n=min_coarse_n, max_levels=max_levels, max_iter=4000, precision="mixed")
    assert mixed["converged"] and mixed["levels"] == want.levels
    assert rel_l2(mixed["u"], want.u) <= 1e-8



def test_pointer_tables_must_match_the_problem_count(ctx):
    """A short pointer table is refused on the host, and a null entry in a
    network table before any launch."""
    a = np.full((8, 8), 2.0)
    pa, pu = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    try:
        ctx.upload(pa, a)
        with pytest.raises(ValueError, match="per problem"):
            ctx.solve_batch_device([8, 8], [pa], [pu, pu])
        with pytest.raises(ValueError, match="per problem"):
            ctx.solve_batch_device([8], [pa], [pu], u0_ptrs=[pu, pu])
        net = capi.Net(ctx, 64, 2)
        try:
            with pytest.raises(ValueError, match="null pointer"):
                net.predict_device([pa, 0], [1.0, 1.0], [pu, pu])
        finally:
            net.close()
    finally:
        ctx.free(pa)
        ctx.free(pu)
