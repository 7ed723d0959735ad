"""Random media generated on the GPU (darcy_grf_fields) against the host
restatement of the reference's sampler (python/darcy/fields.py:31-83 of the
reference) and its committed SciPy fixture.

The GPU sums the orthonormal DCT-III directly in FP64 where SciPy uses an
FFT, so fields agree to round-off (1e-13 of the field's scale) and the media
(the sign of each value) agree exactly wherever a value is not within
round-off of zero.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")
from darcy import capi  # noqa: E402
from darcy.fields import gaussian_random_field, two_phase_permeability  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    with capi.Context() as c:
        yield c


def test_fields_match_the_committed_fixture(ctx):
    fx = np.load(os.path.join(ROOT, "tests", "golden", "grf_seed0.npz"))
    for n in (16, 64):
        for seed in range(3):
            xi = np.random.default_rng(seed).standard_normal((n, n))
            g, a = ctx.grf(xi)
            want = fx[f"grf_n{n}_s{seed}"]
            np.testing.assert_allclose(g[0], want, rtol=0, atol=1e-13 * np.abs(want).max())
            assert np.array_equal(np.packbits(g[0] >= 0.0), fx[f"mask_n{n}_s{seed}"])
            assert np.array_equal(a[0], two_phase_permeability(want))


@pytest.mark.parametrize("n,count", [(256, 24), (97, 5), (1, 3), (130, 2)])
def test_batch_matches_the_host_sampler(ctx, n, count):
    """One generator stream for all samples, as `darcy dataset` draws it."""
    rng = np.random.default_rng(123)
    xi = rng.standard_normal((count, n, n))
    g, a = ctx.grf(xi, alpha=2.0, tau=3.0, hi=12.0, lo=3.0)
    rng = np.random.default_rng(123)
    for i in range(count):
        want = gaussian_random_field(n, seed=rng)
        scale = max(np.abs(want).max(), 1e-300)
        np.testing.assert_allclose(g[i], want, rtol=0, atol=1e-13 * scale)
        clear = np.abs(want) > 1e-12 * scale
        assert np.array_equal(a[i][clear], two_phase_permeability(want)[clear])


def test_other_spectra_and_phases(ctx):
    xi = np.random.default_rng(5).standard_normal((2, 48, 48))
    g, a = ctx.grf(xi, alpha=3.5, tau=7.0, hi=1e3, lo=1e-3)
    # The same draws through the host sampler, sample by sample.
    rng = np.random.default_rng(5)
    for i in range(2):
        want = gaussian_random_field(48, alpha=3.5, tau=7.0, seed=rng)
        np.testing.assert_allclose(g[i], want, rtol=0, atol=1e-13 * np.abs(want).max())
        assert set(np.unique(a[i])) <= {1e3, 1e-3}


def test_device_pointers_and_bad_arguments(ctx):
    n, count = 64, 3
    xi = np.random.default_rng(9).standard_normal((count, n, n))
    px, pa = ctx.alloc(xi.nbytes), ctx.alloc(xi.nbytes)
    try:
        ctx.upload(px, xi)
        ctx.grf_device(count, n, px, pa)
        a = np.empty_like(xi)
        ctx.download(a, pa)
    finally:
        ctx.free(px)
        ctx.free(pa)
    assert np.array_equal(a, ctx.grf(xi, field=False)[1])
    with pytest.raises(ValueError, match="positive"):
        ctx.grf(xi, lo=-1.0)
    with pytest.raises(ValueError):
        ctx.grf(xi, tau=0.0)
