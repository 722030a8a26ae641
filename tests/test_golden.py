"""The committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from the pinned oracle) against the oracle and
the product's host assembly: any change of the operator or of the solver
semantics shows up here."""

import numpy as np
import pytest

from cup2d_amd import poisson as PS

import golden_fixtures as G


def test_pinv_fixture(oracle):
    P = np.load(G.GOLDEN + "/pinv.npy", allow_pickle=False)
    assert P.shape == (64, 64)
    assert np.array_equal(oracle.pinv(), P)


@pytest.mark.parametrize("name", G.CASES)
def test_operator_fixture(oracle, name):
    m, d = G.load(name)
    A = oracle.assemble(m)
    rp, col, val, mr = PS.assemble_csr(m, 0)
    if "row_ptr" in d:
        assert np.array_equal(A.row_ptr, d["row_ptr"])
        assert np.array_equal(A.col, d["col"])
        assert np.array_equal(A.val, d["val"])
    else:
        assert G.csr_digest(A.row_ptr, A.col, A.val) == G.digest(name)
        assert A.nnz == int(d["nnz"])
    # the product's host assembly gives the same bits
    assert G.csr_digest(rp, col, val) == G.csr_digest(A.row_ptr, A.col, A.val)


@pytest.mark.parametrize("name", G.CASES)
def test_solution_fixture(oracle, name):
    m, d = G.load(name)
    A = oracle.assemble(m)
    tr = oracle.true_residual(A, d["b"], d["x"])
    assert tr["rel2"] <= 1e-13
    h2 = np.repeat(oracle.h2(m), 64)
    assert abs(np.dot(h2, d["x"])) <= 1e-13 * np.abs(d["x"]).max() * h2.sum()
    # the oracle reproduces its own fixture
    x, _ = oracle.bicgstab(A, d["b"], rel_tol=1e-14, norm="l2", max_iter=20000)
    x = oracle.mean_remove(oracle.h2(m), x)
    assert np.linalg.norm(x - d["x"]) <= 1e-10 * np.linalg.norm(d["x"])


def test_mean_row_fixture(oracle):
    m, d = G.load("uniform16x8")
    A1 = oracle.assemble(m, 1)
    assert A1.mean_row == int(d["mean_row"])
    # both mean treatments give the same mean-free pressure
    x0, x1 = d["x"], d["x_mean_row"]
    assert np.linalg.norm(x0 - x1) <= 1e-8 * np.linalg.norm(x0)


def test_cylinder_record_fixture():
    """The CPU run's record of the 100-step cylinder flow (make_cylinder100.py):
    its shape, and its first time step, which from rest is calcMaxTimestep's
    closed form with umax = |uinf| (Simulation.cpp:426-460)."""
    from cup2d_amd import workloads as W

    rec = np.load(G.GOLDEN + "/cylinder100.npz", allow_pickle=False)
    m = W.cylinder_mesh()
    assert int(rec["n_blocks"]) == m.n_blocks
    assert rec["dt"].shape == (100,) and rec["iterations"].shape == (100,)
    assert (rec["dt"] > 0).all() and (rec["iterations"] > 0).all()
    for k in (10, 100):
        for f in "uvp":
            assert rec[f"{f}_mean_{k}"].shape == (m.n_blocks,)
            assert rec[f"{f}_absmax_{k}"].shape == (m.n_blocks,)
    h = float(m.cell_h().min())
    umax = W.CYL_U
    dt0 = min(0.25 * h * h / (W.CYL_NU + 0.25 * h * umax), W.CYL_CFL * (h / (umax + 1e-8)))
    assert rec["dt"][0] == dt0
