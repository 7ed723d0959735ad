"""The command line's host side: the reference's defaults, and errors that
stop before any solve (python/darcy/cli.py:16-205)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

cli = pytest.importorskip("darcy.cli")


def test_defaults_are_the_references():
    p = cli.build_parser()
    s = p.parse_args(["solve"])
    assert (s.grid, s.seed, s.alpha, s.tau, s.f, s.key, s.input, s.output) == (128, 0, 2.0, 3.0, 1.0, "a", None, None)
    d = p.parse_args(["dataset"])
    assert (d.n_samples, d.grid, d.seed, d.alpha, d.tau, d.f, d.threads, d.dtype) == (
        1000, 64, 0, 2.0, 3.0, 1.0, 0, "float32")
    assert d.output == Path("data/darcy_train.npz")
    for a in (s, d):
        assert (a.tol, a.max_iter, a.preconditioner, a.pre_sweeps, a.post_sweeps, a.cycle_index) == (
            1e-10, 10_000, "multigrid", 1, 1, 2)
        assert a.precision == "fp64"


def test_bad_choices_exit_2():
    with pytest.raises(SystemExit) as e:
        cli.build_parser().parse_args(["solve", "--preconditioner", "ilu"])
    assert e.value.code == 2


def test_solve_rejects_a_3d_field_before_solving(tmp_path, capsys):
    f = tmp_path / "a.npy"
    np.save(f, np.ones((2, 4, 4)))
    assert cli.main(["solve", "--input", str(f)]) == 2
    assert "expected a 2-D field" in capsys.readouterr().err


def test_info_prints_build_details(capsys):
    assert cli.main(["info"]) == 0
    out = capsys.readouterr().out
    assert '"version"' in out and '"device"' in out


def test_bench_defaults_are_the_references():
    b = cli.build_parser().parse_args(["bench"])
    assert (b.grids, b.repeats, b.no_scipy, b.preconditioners, b.json) == (
        [64, 128, 256, 512], 3, False, ["multigrid", "jacobi"], None)


@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_direct_operator_is_the_oracles(n, oracle):
    """The benchmark's assembled matrix applies the solver's operator."""
    direct = pytest.importorskip("darcy.direct")
    from conftest import two_phase

    a = two_phase(n, seed=n)
    x = np.random.default_rng(n).standard_normal((n, n))
    got = (direct.assemble(a) @ x.ravel()).reshape(n, n)
    np.testing.assert_allclose(got, oracle.apply_operator(a, x), rtol=1e-13, atol=1e-13 * np.abs(got).max())


def test_direct_solution_is_the_oracles(oracle):
    direct = pytest.importorskip("darcy.direct")
    from conftest import rel_l2, two_phase

    a = two_phase(48, seed=2)
    assert rel_l2(direct.solve_direct(a), oracle.solve(a).u) <= 1e-9
