"""GPU solves of the golden fixtures (single and batched, FP64 and mixed)
against the committed oracle solutions: relative L2 error <= 1e-8 after the
same h^2-weighted mean removal, at 1e-12 relative residual."""

import numpy as np
import pytest

from cup2d_amd import poisson as PS

import golden_fixtures as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", G.CASES)
@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_single_solve_matches_fixture(name, precision):
    m, d = G.load(name)
    s = PS.Solver(precision=precision, norm=PS.NORM_L2, max_iter=4000).set_mesh(m)
    x, st = s.solve(d["b"], rel_tol=1e-12)
    assert st["converged"], st
    assert np.linalg.norm(x - d["x"]) <= 1e-8 * np.linalg.norm(d["x"])


def test_single_solve_mean_row_matches_fixture():
    m, d = G.load("uniform16x8")
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000, mean_constraint=1).set_mesh(m)
    x, st = s.solve(d["b"], rel_tol=1e-12)
    assert st["converged"], st
    assert np.linalg.norm(x - d["x_mean_row"]) <= 1e-8 * np.linalg.norm(d["x_mean_row"])


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_matches_fixtures(precision):
    cases = [G.load(n) for n in G.CASES]
    bs = PS.BatchSolver(PS.default_config(precision=precision, norm=PS.NORM_L2,
                                          max_iter=4000)).set_meshes([m for m, _ in cases])
    x, st = bs.solve(np.concatenate([d["b"] for _, d in cases]), rel_tol=1e-12)
    for k, (_, d) in enumerate(cases):
        assert st[k]["converged"], st[k]
        xk = bs.split(x)[k]
        assert np.linalg.norm(xk - d["x"]) <= 1e-8 * np.linalg.norm(d["x"])
