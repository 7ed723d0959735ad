"""darcy_solve_batch_noalloc (the library's static scratch array, no device
allocation) against darcy_solve_batch on the same device-resident inputs:
the same bits and statistics, on calls in a row (a batch with a side of 384,
which runs on 1024-thread workgroups, then a smaller batch), and its size
and pointer checks."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import two_phase

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")
from darcy import capi  # noqa: E402


def _device_batch(ctx, media, u0s=None):
    bufs = []
    for i, a in enumerate(media):
        pa, pu = ctx.alloc(a.nbytes), ctx.alloc(a.nbytes)
        ctx.upload(pa, a)
        p0 = None
        if u0s is not None:
            p0 = ctx.alloc(a.nbytes)
            ctx.upload(p0, u0s[i])
        bufs.append((pa, pu, p0))
    return bufs


def _solve(ctx, media, bufs, no_alloc, **kw):
    sizes = [a.shape[0] for a in media]
    u0 = [b[2] for b in bufs] if bufs[0][2] is not None else None
    print(f"{'noalloc' if no_alloc else 'batch'} sizes={sizes} {kw}", flush=True)
    res = ctx.solve_batch_device(sizes, [b[0] for b in bufs], [b[1] for b in bufs], u0_ptrs=u0, no_alloc=no_alloc,
                                 **kw)
    us = []
    for (_, pu, _), a in zip(bufs, media):
        u = np.empty_like(a)
        ctx.download(u, pu)
        us.append(u)
    return res, us


def _free(ctx, bufs):
    for b in bufs:
        for p in b:
            if p:
                ctx.free(p)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_same_bits_as_the_allocating_entry_point_twice(precision):
    rng = np.random.default_rng(4)
    with capi.Context() as ctx:
        assert ctx.scratch_bytes() == 64 << 20
        first = [two_phase(n, seed=50 + i) for i, n in enumerate([64, 33, 128, 17, 96, 5, 256])]
        u0s = [1e-3 * rng.standard_normal(a.shape) for a in first]
        # sides >= batch_wide_side (384): both entry points on 1024-thread workgroups
        wide = [two_phase(n, seed=60 + i) for i, n in enumerate([384, 40])]
        second = [two_phase(n, seed=70 + i) for i, n in enumerate([16, 8, 24])]
        for media, warm in ((first, u0s), (wide, None), (second, None)):
            bufs = _device_batch(ctx, media, warm)
            try:
                ra, ua = _solve(ctx, media, bufs, False, precision=precision)
                rb, ub = _solve(ctx, media, bufs, True, precision=precision)
            finally:
                _free(ctx, bufs)
            assert rb["converged"].all()
            for key in ("iterations", "converged", "residual", "levels"):
                assert np.array_equal(ra[key], rb[key]), key
            for x, y in zip(ua, ub):
                assert np.array_equal(x, y)


def test_rejects_what_does_not_fit_and_host_pointers():
    with capi.Context() as ctx:
        a = two_phase(2048, seed=1)
        bufs = _device_batch(ctx, [a])
        try:
            with pytest.raises(ValueError, match="scratch"):
                _solve(ctx, [a], bufs, True)
        finally:
            _free(ctx, bufs)
        o = ctx._opts()  # host pointers
        small = two_phase(8, seed=2)
        out = np.empty_like(small)
        n = (C.c_int64 * 1)(8)
        ap = (C.c_void_p * 1)(small.ctypes.data)
        up = (C.c_void_p * 1)(out.ctypes.data)
        st = (capi.Stats * 1)()
        rc = ctx.lib.darcy_solve_batch_noalloc(ctx.ptr, 1, n, ap, None, None, C.byref(o), None, up, st)
        assert rc == capi.INVALID_INPUT
        assert "device pointers" in ctx.lib.darcy_last_error(ctx.ptr).decode()
        # still usable afterwards
        assert ctx.solve(two_phase(16, seed=3))["converged"]
