"""C ABI and host-side behaviour that needs no GPU.

* libdarcy_hip.so loads and exports every function include/darcy_hip.h
  declares (no compute call is made).
* The pybind11 surface rejects bad arguments with the reference's
  exceptions and messages before any device work (reference
  tests/test_api.py:54-92, src/bindings.cpp:33-78).
* The GRF inputs reproduce the committed fixture (SciPy's idctn path) and
  the numpy-only DCT-III agrees with SciPy.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import darcy
from darcy import capi
from darcy.fields import dct3_ortho, gaussian_random_field, two_phase_permeability

HEADER = os.path.join(ROOT, "include", "darcy_hip.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(darcy_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions()
    assert "darcy_solve" in names and "darcy_solve_batch" in names
    lib = ctypes.CDLL(capi._LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_options_are_the_reference_defaults():
    lib = capi._load()
    o = capi.Opts()
    lib.darcy_default_opts(ctypes.byref(o))
    assert (o.tol, o.max_iter, o.preconditioner, o.pre_sweeps, o.post_sweeps, o.cycle_index) == (
        1e-10, 10000, 2, 1, 1, 2)
    assert (o.record_history, o.precision, o.device_pointers) == (0, 0, 0)
    assert (o.min_coarse_n, o.max_levels, o.breakdown_at) == (1, 32, 0)  # multigrid.hpp:43-46; no injection


def test_num_levels_closed_form():
    lib = capi._load()
    lib.darcy_num_levels.restype = ctypes.c_int32
    lib.darcy_num_levels.argtypes = [ctypes.c_int64]
    for n in [1, 2, 3, 4, 5, 8, 9, 64, 96, 192, 256, 384, 512, 1024, 4096]:
        expected = 1 if n == 1 else int(np.ceil(np.log2(n))) + 1
        assert lib.darcy_num_levels(n) == expected == darcy.num_levels(n)


class TestValidationBeforeDeviceWork:
    def test_rejects_non_square(self):
        with pytest.raises(ValueError, match="square"):
            darcy.solve(np.ones((4, 8)))

    def test_rejects_wrong_dimensionality(self):
        with pytest.raises(ValueError):
            darcy.solve(np.ones((4, 4, 4)))

    def test_rejects_empty(self):
        with pytest.raises(ValueError):
            darcy.solve(np.ones((0, 0)))

    def test_rejects_mismatched_forcing_shape(self):
        with pytest.raises(ValueError, match="same shape"):
            darcy.solve(np.ones((8, 8)), np.ones((4, 4)))

    def test_rejects_unknown_preconditioner(self):
        with pytest.raises(ValueError, match="unknown preconditioner"):
            darcy.solve(np.ones((8, 8)), preconditioner="ilu")

    @pytest.mark.parametrize("kwargs", [{"tol": 0.0}, {"tol": -1e-3}, {"max_iter": 0}])
    def test_rejects_bad_solver_parameters(self, kwargs):
        with pytest.raises(ValueError):
            darcy.solve(np.ones((8, 8)), **kwargs)

    @pytest.mark.parametrize("kwargs", [{"pre_sweeps": 0}, {"post_sweeps": 0}, {"cycle_index": 3}])
    def test_rejects_bad_multigrid_parameters(self, kwargs):
        with pytest.raises(ValueError):
            darcy.solve(np.ones((8, 8)), **kwargs)

    def test_rejects_unknown_precision(self):
        with pytest.raises(ValueError, match="precision"):
            darcy.solve(np.ones((8, 8)), precision="fp16")

    def test_batch_rejects_wrong_shape(self):
        with pytest.raises(ValueError, match=r"batch, n, n"):
            darcy.solve_batch(np.ones((8, 8)))

    def test_apply_operator_checks_shapes_and_values(self):
        with pytest.raises(ValueError, match="same shape"):
            darcy.apply_operator(np.ones((8, 8)), np.ones((4, 4)))
        a = np.full((8, 8), 2.0)
        a[3, 4] = -1.0
        with pytest.raises(ValueError, match="positive"):
            darcy.apply_operator(a, np.ones((8, 8)))

    def test_invalid_input_is_a_value_error(self):
        assert issubclass(darcy._core.InvalidInput, ValueError)


class TestIntrospection:
    def test_build_info_has_the_reference_keys(self):
        info = darcy.build_info()
        assert set(info) == {"version", "compiler", "openmp", "max_threads", "numpy"}
        assert isinstance(info["openmp"], bool) and info["max_threads"] >= 1

    def test_version(self):
        assert darcy.__version__.count(".") == 2

    def test_thread_count_round_trips(self):
        original = darcy.num_threads()
        try:
            darcy.set_num_threads(2)
            assert darcy.num_threads() == (2 if darcy.has_openmp() else 1)
        finally:
            darcy.set_num_threads(original)


class TestFields:
    def test_matches_committed_fixture(self):
        fx = np.load(os.path.join(ROOT, "tests", "golden", "grf_seed0.npz"))
        for n in (16, 64):
            for seed in range(3):
                g = gaussian_random_field(n, seed=seed)
                np.testing.assert_allclose(g, fx[f"grf_n{n}_s{seed}"], rtol=0, atol=1e-14)
                assert np.array_equal(np.packbits(g >= 0.0), fx[f"mask_n{n}_s{seed}"])

    def test_numpy_dct3_matches_scipy(self):
        from scipy.fft import idctn

        x = np.random.default_rng(3).standard_normal((48, 48))
        np.testing.assert_allclose(dct3_ortho(x), idctn(x, norm="ortho"), atol=1e-13)

    def test_dct3_closed_form(self):
        """DCT-III of a unit impulse at k = 0 is the constant 1/sqrt(N) per axis."""
        n = 12
        x = np.zeros((n, n))
        x[0, 0] = 1.0
        np.testing.assert_allclose(dct3_ortho(x), np.full((n, n), 1.0 / n), atol=1e-15)

    def test_reproducible_mean_zero_two_valued(self):
        assert np.array_equal(gaussian_random_field(32, seed=42), gaussian_random_field(32, seed=42))
        assert not np.array_equal(gaussian_random_field(32, seed=1), gaussian_random_field(32, seed=2))
        assert abs(gaussian_random_field(128, seed=0).mean()) < 1e-10
        a = two_phase_permeability(gaussian_random_field(64, seed=0), hi=9.0, lo=2.0)
        assert set(np.unique(a)) == {2.0, 9.0}

    def test_rejects_invalid_parameters(self):
        with pytest.raises(ValueError):
            gaussian_random_field(0)
        with pytest.raises(ValueError, match="positive"):
            two_phase_permeability(np.zeros((8, 8)), hi=1.0, lo=-1.0)


def test_one_draw_equals_the_per_sample_draws():
    """`darcy dataset` draws all normals of a chunk at once; numpy's
    generator yields the same sequence as the reference's per-sample
    (n, n) draws (python/darcy/cli.py:79-84 of the reference)."""
    a = np.random.default_rng(7).standard_normal((5, 24, 24))
    rng = np.random.default_rng(7)
    b = np.stack([rng.standard_normal((24, 24)) for _ in range(5)])
    assert np.array_equal(a, b)
