"""Launch-shape knobs: HIP-graph replay of the grid path's cycle and the
autotuner.  No knob may change a result, so every variant is compared bit
for bit with the eager launch."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import two_phase

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("darcy.capi")
autotune = pytest.importorskip("darcy.autotune")


def _solve(ctx, a, precision, cycle_index=2):
    n = a.shape[0]
    pa, pu = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    try:
        ctx.upload(pa, a)
        st = ctx.solve_device(n, pa, pu, precision=precision, cycle_index=cycle_index)
        u = np.empty_like(a)
        ctx.download(u, pu)
    finally:
        ctx.free(pa)
        ctx.free(pu)
    return u, st["iterations"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_eager_launches(precision):
    with capi.Context(apply_tuning=False) as ctx:
        ctx.set_tuning("grid_min_n", 64)
        ctx.set_tuning("grid_tail_side", 64)
        a300, a200 = two_phase(300, seed=3), two_phase(200, seed=4)
        print(f"graph replay {precision}: eager references", flush=True)
        ctx.set_tuning("grid_graph", 0)
        u300, it300 = _solve(ctx, a300, precision)
        u200, it200 = _solve(ctx, a200, precision)
        ctx.set_tuning("grid_graph", 1)
        # capture, reuse, recapture for another size, capture again: every
        # replay interleaved with the eager CG kernels and host syncs
        for a, u_ref, it_ref in ((a300, u300, it300), (a300, u300, it300), (a200, u200, it200),
                                 (a300, u300, it300)):
            print(f"graph replay {precision}: n={a.shape[0]}", flush=True)
            u, it = _solve(ctx, a, precision)
            assert it == it_ref
            assert np.array_equal(u, u_ref)


def test_autotuner_picks_valid_shapes_and_restores_the_context():
    with capi.Context(apply_tuning=False) as ctx:
        before = {k: ctx.get_tuning(k) for k in autotune.KEYS}
        print("autotune: small run", flush=True)
        table = autotune.tune(ctx, sizes=[128, 200], small=[64, 96], repeats=1, log=lambda s: None, batch_n=64,
                              batch_count=32)
        after = {k: ctx.get_tuning(k) for k in autotune.KEYS}
        assert after == before
        chosen = table["chosen"]
        assert chosen["grid_tail_side"] in (32, 64, 128) and chosen["grid_graph"] in (0, 1)
        assert chosen["grid_min_n"] in (63, 64, 96)  # bounded below by the smallest size minus one
        assert chosen["batch_step_cells"] in autotune.STEP_CELLS
        assert chosen["batch_step_cells_fp32"] in autotune.STEP_CELLS
        assert len(table["results"]["grid"]) == 6 and len(table["results"]["switch"]) == 2
        assert len(table["results"]["batch"]) == 2 * len(autotune.STEP_CELLS)
        autotune.apply(ctx, table)
        assert {k: ctx.get_tuning(k) for k in autotune.KEYS} == chosen


def _solve_persistent(ctx, a, precision):
    n = a.shape[0]
    pa, pu = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
    try:
        ctx.upload(pa, a)
        st = ctx.solve_batch_device([n], [pa], [pu], precision=precision)
        u = np.empty_like(a)
        ctx.download(u, pu)
    finally:
        ctx.free(pa)
        ctx.free(pu)
    return u, int(st["iterations"][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tail_side_equals_persistent_kernel(precision):
    """n = 300 has sides 300, 150, 75, 38, 19, 10, 5, 3, 2, 1.  Tail sides
    32, 64 and 128 start the single-workgroup tail at 19, 38 and 75: its
    first level then sits in LDS (staged in and out, with odd lengths) or
    in global memory, depending on the precision's LDS limit (64 in FP32,
    32 in FP64).  Every variant gives the persistent kernel's bits."""
    a = two_phase(300, seed=11)
    with capi.Context(apply_tuning=False) as ctx:
        ctx.set_tuning("grid_min_n", 64)
        ctx.set_tuning("grid_fp64_dense_tail", 0)  # the recursive cycle: the persistent kernel's bits
        ctx.set_tuning("grid_dense_side", 8)  # the FP32 cycle's tail as the persistent kernel runs it
        u_ref, it_ref = _solve_persistent(ctx, a, precision)
        for side in (32, 64, 128):
            for graph in (0, 1):
                print(f"tail side {side} graph {graph} {precision}", flush=True)
                ctx.set_tuning("grid_tail_side", side)
                ctx.set_tuning("grid_graph", graph)
                u, it = _solve(ctx, a, precision)
                assert it == it_ref
                assert np.array_equal(u, u_ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_step_cells_do_not_change_results(precision):
    """The persistent kernel's pipeline step (its sweep tile) over an uneven
    batch whose sides put 1 to 32 rows in a step: the same bits for every
    step size."""
    sizes = [37, 64, 100, 129, 256]
    media = [two_phase(n, seed=60 + i) for i, n in enumerate(sizes)]
    with capi.Context(apply_tuning=False) as ctx:
        bufs = [(ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)) for a in media]
        try:
            for (pa, _), a in zip(bufs, media):
                ctx.upload(pa, a)
            ref = None
            for cells in (128, 1024, 4096, 16384):
                print(f"batch step {cells} {precision}", flush=True)
                ctx.set_tuning("batch_step_cells", cells)
                ctx.set_tuning("batch_step_cells_fp32", cells)
                st = ctx.solve_batch_device(sizes, [b[0] for b in bufs], [b[1] for b in bufs], precision=precision)
                us = []
                for (_, pu), a in zip(bufs, media):
                    u = np.empty_like(a)
                    ctx.download(u, pu)
                    us.append(u)
                if ref is None:
                    ref = (st["iterations"].copy(), us)
                    assert st["converged"].all()
                else:
                    assert np.array_equal(st["iterations"], ref[0])
                    for u, u_ref in zip(us, ref[1]):
                        assert np.array_equal(u, u_ref)
        finally:
            for pa, pu in bufs:
                ctx.free(pa)
                ctx.free(pu)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_workgroups_do_not_change_results(precision, oracle):
    """A batch whose largest side is >= batch_wide_side runs on 1024-thread
    workgroups (one per CU): an uneven batch with sides 1 ... 512 gives the
    same bits and iteration counts on narrow workgroups (0, 1024) and wide
    ones (384, 64); the 512^2 problem's u is within 1e-9 (FP64) / 1e-8 (FP32
    cycle, against the oracle's FP32 restatement) of the oracle's, its
    iterations equal."""
    from conftest import rel_l2

    sizes = [512, 37, 384, 64, 1, 100, 2, 256, 3, 129]
    media = [two_phase(n, seed=80 + i) for i, n in enumerate(sizes)]
    with capi.Context(apply_tuning=False) as ctx:
        bufs = [(ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)) for a in media]
        try:
            for (pa, _), a in zip(bufs, media):
                ctx.upload(pa, a)
            ref = None
            for side in (0, 384, 1024, 64):
                print(f"batch wide side {side} {precision}", flush=True)
                ctx.set_tuning("batch_wide_side", side)
                assert ctx.get_tuning("batch_wide_side") == side
                st = ctx.solve_batch_device(sizes, [b[0] for b in bufs], [b[1] for b in bufs], precision=precision)
                us = []
                for (_, pu), a in zip(bufs, media):
                    u = np.empty_like(a)
                    ctx.download(u, pu)
                    us.append(u)
                if ref is None:
                    ref = (st["iterations"].copy(), us)
                    assert st["converged"].all()
                else:
                    assert np.array_equal(st["iterations"], ref[0])
                    for u, u_ref in zip(us, ref[1]):
                        assert np.array_equal(u, u_ref)
        finally:
            for pa, pu in bufs:
                ctx.free(pa)
                ctx.free(pu)
    want = oracle.solve(media[0], fp32=precision == "mixed")
    assert int(ref[0][0]) == want.iterations
    assert rel_l2(ref[1][0], want.u) <= (1e-9 if precision == "fp64" else 1e-8)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_workgroups_name_the_first_bad_problem(precision):
    """A batch on 1024-thread workgroups (largest side 400) with invalid
    media: the kernel's validation names the lowest bad problem and its
    first offending flat index, as on narrow workgroups (solver.hpp:68-77)."""
    media = [two_phase(n, seed=90 + i) for i, n in enumerate([16, 400, 40])]
    media[1].flat[1234] = -1.0
    media[1].flat[5000] = 0.0
    media[2][1, 1] = np.nan
    with capi.Context(apply_tuning=False) as ctx:
        assert ctx.get_tuning("batch_wide_side") == 384
        bufs = [(ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)) for a in media]
        try:
            for (pa, _), a in zip(bufs, media):
                ctx.upload(pa, a)
            print(f"wide batch with invalid media {precision}", flush=True)
            with pytest.raises(ValueError, match=r"flat index 1234 of problem 1"):
                ctx.solve_batch_device([16, 400, 40], [b[0] for b in bufs], [b[1] for b in bufs], precision=precision)
        finally:
            for pa, pu in bufs:
                ctx.free(pa)
                ctx.free(pu)


@pytest.mark.parametrize("n", [130, 300, 1024])
def test_fp64_dense_tail_matches_the_recursive_cycle(n, oracle):
    """The grid path's FP64 dense tail applies the same linear map as the
    recursive cycle over the levels of side <= 8, rounded differently: the
    same iteration count and solutions within 1e-12 of each other (and of
    the oracle's to 1e-9), with and without graph replay."""
    from conftest import rel_l2

    a = two_phase(n, seed=21)
    with capi.Context(apply_tuning=False) as ctx:
        ctx.set_tuning("grid_min_n", 64)
        ctx.set_tuning("grid_fp64_dense_tail", 0)
        u_rec, it_rec = _solve(ctx, a, "fp64")
        for graph in (0, 1):
            print(f"dense tail n={n} graph {graph}", flush=True)
            ctx.set_tuning("grid_fp64_dense_tail", 1)
            ctx.set_tuning("grid_graph", graph)
            u, it = _solve(ctx, a, "fp64")
            assert it == it_rec
            assert rel_l2(u, u_rec) <= 1e-12
        assert ctx.get_tuning("grid_fp64_dense_tail") == 1
    if n <= 300:
        assert rel_l2(u, oracle.solve(a).u) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [130, 300, 1024])
def test_dense_side_16_matches_side_8(precision, n, oracle):
    """The grid path's tail from the first level of side <= 16 down as one
    dense matrix read from L2 (grid_dense_side = 16, the default; n = 130 and
    300 reach that level at sides 9 and 10, 1024 at 16; with the W-cycle the
    matrix of both visits of that level, 2M - M A M, applied once per pair)
    applies the same linear map as the dense tail over the levels of side
    <= 8, rounded differently: iteration counts within one, solutions within
    1e-10 of each other and within 1e-9 (FP64) / 1e-8 (FP32 cycle) of the
    oracle's; graph replay gives the eager launches' bits."""
    from conftest import rel_l2

    a = two_phase(n, seed=23)
    with capi.Context(apply_tuning=False) as ctx:
        ctx.set_tuning("grid_min_n", 64)
        ctx.set_tuning("grid_dense_side", 8)
        u8, it8 = _solve(ctx, a, precision)
        ctx.set_tuning("grid_dense_side", 16)
        runs = []
        for graph in (0, 1):
            print(f"dense side 16 n={n} {precision} graph {graph}", flush=True)
            ctx.set_tuning("grid_graph", graph)
            runs.append(_solve(ctx, a, precision))
        assert ctx.get_tuning("grid_dense_side") == 16
    (u_e, it_e), (u_g, it_g) = runs
    assert it_e == it_g and np.array_equal(u_e, u_g)
    assert abs(it_e - it8) <= 1
    assert rel_l2(u_e, u8) <= 1e-10
    if n <= 300:
        assert rel_l2(u_e, oracle.solve(a).u) <= (1e-9 if precision == "fp64" else 1e-8)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_side_16_v_cycle(precision, oracle):
    """With the V-cycle the side-16 matrix is the single visit's (one visit
    per coarse level, no pair): n = 300, same bar as the W-cycle test, and
    the oracle's V-cycle iteration count within one."""
    from conftest import rel_l2

    a = two_phase(300, seed=29)
    with capi.Context(apply_tuning=False) as ctx:
        ctx.set_tuning("grid_min_n", 64)
        ctx.set_tuning("grid_dense_side", 8)
        u8, it8 = _solve(ctx, a, precision, cycle_index=1)
        ctx.set_tuning("grid_dense_side", 16)
        print(f"dense side 16 V-cycle n=300 {precision}", flush=True)
        u16, it16 = _solve(ctx, a, precision, cycle_index=1)
    ref = oracle.solve(a, cycle_index=1)
    assert abs(it16 - it8) <= 1 and abs(it16 - ref.iterations) <= 1
    assert rel_l2(u16, u8) <= 1e-10
    assert rel_l2(u16, ref.u) <= (1e-9 if precision == "fp64" else 1e-8)
