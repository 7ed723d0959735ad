"""The command line on the GPU: ``dataset`` draws the reference's media
stream (one default_rng(seed) for all samples) and stores solutions that the
oracle reproduces; ``solve`` reports and writes one problem."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

darcy = pytest.importorskip("darcy")
cli = pytest.importorskip("darcy.cli")


def test_dataset_media_and_solutions(tmp_path, oracle, capsys):
    out = tmp_path / "d.npz"
    print("cli dataset 6 x 32^2", flush=True)
    assert cli.main(["dataset", "--n-samples", "6", "--grid", "32", "--seed", "3", "--output", str(out)]) == 0
    assert "wrote" in capsys.readouterr().out
    with np.load(out) as z:
        a, u = z["a"], z["u"]
    assert a.dtype == np.float32 and u.dtype == np.float32 and a.shape == u.shape == (6, 32, 32)
    rng = np.random.default_rng(3)
    for i in range(6):
        ai = darcy.two_phase_permeability(darcy.gaussian_random_field(32, seed=rng))
        assert np.array_equal(a[i], ai.astype(np.float32))
        want = oracle.solve(ai)
        assert rel_l2(u[i].astype(np.float64), want.u) <= 1e-6  # stored as float32


def test_solve_report_and_output(tmp_path, capsys):
    out = tmp_path / "s.npz"
    print("cli solve 64^2 seed 0", flush=True)
    assert cli.main(["solve", "--grid", "64", "--seed", "0", "--output", str(out)]) == 0
    text = capsys.readouterr().out
    assert "iterations    21" in text and "(7 levels)" in text and "converged     True" in text
    with np.load(out) as z:
        r = darcy.solve(z["a"])
        assert np.array_equal(z["u"], r.u)


def test_solve_mixed_and_nonconvergence(tmp_path):
    print("cli solve mixed / max-iter 2", flush=True)
    assert cli.main(["solve", "--grid", "96", "--precision", "mixed"]) == 0
    assert cli.main(["solve", "--grid", "64", "--max-iter", "2"]) == 1


def test_bench_rows(tmp_path, capsys):
    out = tmp_path / "b.json"
    print("cli bench 32, 64", flush=True)
    assert cli.main(["bench", "--grids", "32", "64", "--repeats", "1", "--json", str(out)]) == 0
    assert "| grid | preconditioner | CG iters |" in capsys.readouterr().out
    import json

    rows = json.loads(out.read_text())
    mg = {r["grid"]: r for r in rows if r["preconditioner"] == "multigrid"}
    assert (mg[32]["iterations"], mg[32]["levels"], mg[64]["iterations"], mg[64]["levels"]) == (18, 6, 21, 7)
    # Both stop at a 1e-10 relative residual; the error is up to the condition
    # number (~ n^2 x contrast) times that, which multigrid CG's iterates do
    # not approach but Jacobi CG's may.
    assert all(r["relative_error"] <= 1e-9 for r in mg.values())
    assert all(r["relative_error"] <= 1e-5 for r in rows)
