"""The large and batched workloads at their real sizes.

* 4096^2 single solves, FP64 and the FP32 cycle, on the two-phase (12/3)
  and the 1e6-contrast (1e3/1e-3) media of seed 0: converged on 13 levels,
  the FP32-cycle solution within 1e-9 relative L2 of the FP64 one (two
  different preconditioned iterations reaching the same A^-1 b), and the
  true residual (a vectorised restatement of the operator, computed on the
  host).  On the two-phase medium the true residual is below 1e-9.  At 1e6
  contrast the reference's CG itself lets the recursively updated residual
  drift from b - A u: the oracle (bit-exact with the reference) reports
  4e-11 / 9e-11 at 256^2 / 1024^2 while b - A u is 8e-9 / 1e-8, and the gap
  grows with the condition number and the iteration count (346 at 4096^2),
  so there the bound is the drift's scale, 1e-4.  The CPU oracle takes
  minutes at 4096^2; it bounds the same kernels at 1024^2 in both media
  (tests/test_gpu_grid.py).
* An uneven batch with two problems of every side in {64, 96, 128, 192, 256,
  384, 512} in one persistent launch, each against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2, two_phase
from test_gpu_grid import true_residual

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")

MEDIA = {"two_phase": (12.0, 3.0), "contrast_1e6": (1e3, 1e-3)}


@pytest.fixture(scope="module")
def grf4096():
    from darcy.fields import gaussian_random_field

    return gaussian_random_field(4096, seed=0)


@pytest.mark.parametrize("medium", sorted(MEDIA))
def test_4096_fp64_and_fp32_cycle(grf4096, medium):
    hi, lo = MEDIA[medium]
    a = np.where(grf4096 >= 0.0, hi, lo)
    bound = 1e-9 if medium == "two_phase" else 1e-4
    fp64 = darcy.solve(a, max_iter=2000)
    assert fp64.converged and fp64.levels == 13 and fp64.residual < 1e-10
    assert true_residual(a, fp64.u) < bound
    mixed = darcy.solve(a, max_iter=2000, precision="mixed")
    assert mixed.converged and mixed.levels == 13 and mixed.residual < 1e-10
    assert true_residual(a, mixed.u) < bound
    assert rel_l2(mixed.u, fp64.u) <= 1e-9
    if medium == "two_phase":
        assert fp64.iterations == 35 and mixed.iterations <= 40


SIDES = [64, 96, 128, 192, 256, 384, 512]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uneven_sides_up_to_512_in_one_launch(oracle, precision):
    sizes = [n for n in SIDES for _ in range(2)]
    media = [two_phase(n, seed=1000 + i) for i, n in enumerate(sizes)]
    us, res = darcy.solve_batch_uneven(media, precision=precision)
    assert res.all_converged
    for a, u, it in zip(media, us, res.iterations):
        want = oracle.solve(a, fp32=precision == "mixed")
        assert it == want.iterations, (a.shape, it, want.iterations)
        assert rel_l2(u, want.u) <= 1e-9
