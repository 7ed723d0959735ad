"""Writes the golden fixtures of tests/golden/ (run in a container that has
SciPy and, for the iteration table, oracle/_ref built from the reference's
own headers).

  grf_seed0.npz          GRF values and two-phase masks for seeds 0..2 at
                         n = 16, 64 (scipy.fft.idctn path, the reference's
                         call at python/darcy/fields.py:66-69)
  ref_iterations.json    iteration counts and levels of the reference solver
                         (oracle/_ref) on two-phase and 1e6-contrast media at
                         sizes the README does not publish

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "darcy-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from darcy.fields import gaussian_random_field, two_phase_permeability  # noqa: E402


def grf_fixture() -> None:
    out = {}
    for n in (16, 64):
        for seed in range(3):
            g = gaussian_random_field(n, seed=seed)
            out[f"grf_n{n}_s{seed}"] = g
            out[f"mask_n{n}_s{seed}"] = np.packbits(g >= 0.0)
    np.savez_compressed(os.path.join(HERE, "grf_seed0.npz"), **out)


def iteration_fixture() -> None:
    import pyoracle

    ref = pyoracle.Reference()
    pyoracle.Oracle().set_num_threads(1)  # serial reductions: a reproducible table
    rows = []
    for n in (16, 17, 31, 33, 47, 96, 100, 192, 384):
        for seed in (0, 1):
            for hi, lo, medium in ((12.0, 3.0, "two_phase"), (1e3, 1e-3, "contrast_1e6")):
                a = two_phase_permeability(gaussian_random_field(n, seed=seed), hi=hi, lo=lo)
                for cycle in (1, 2):
                    r = ref.solve(a, cycle_index=cycle, max_iter=5000)
                    rows.append({"n": n, "seed": seed, "medium": medium, "cycle_index": cycle,
                                 "iterations": r.iterations, "levels": r.levels, "converged": r.converged})
    with open(os.path.join(HERE, "ref_iterations.json"), "w") as f:
        json.dump(rows, f, indent=0)


if __name__ == "__main__":
    grf_fixture()
    iteration_fixture()
