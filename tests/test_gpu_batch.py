"""GPU tests of the batched solver (cup2d_psb_*): many independent meshes of
uneven sizes solved together, each instance checked on its own.

Bars: every instance meets the same FP64 true residual as a single solve
(1e-10 relative) and matches the oracle's solution; the FP64 batch agrees
with the single-solve FP64 path instance by instance to 1e-9 (the same
algorithm, partial sums reduced in a different order).
"""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_meshes():
    ms = [W.uniform_16x8_mesh(), M.uniform(4, 2, 2.0)]
    ms += W.batch64_meshes(seed=7, count=3, nmin=64, nmax=600)
    # periodic in both directions, 4 levels: wrap-around same-level and
    # coarse-fine neighbours inside a batch
    ms.append(M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)], [0.3, 0.15, 0.08],
                       True, True))
    return ms


@pytest.fixture(scope="module")
def batch_rhs(batch_meshes):
    return [W.vortex_rhs(m, 20 + k) for k, m in enumerate(batch_meshes)]


@pytest.fixture(scope="module")
def oracle_solutions(oracle, batch_meshes, batch_rhs):
    out = []
    for m, b in zip(batch_meshes, batch_rhs):
        A = oracle.assemble(m)
        x, _ = oracle.bicgstab(A, b, rel_tol=1e-13, norm="l2", max_iter=6000)
        out.append(oracle.mean_remove(oracle.h2(m), x))
    return out


def make_batch(meshes, precision=PS.FP64, item_blocks=0, **cfg):
    c = PS.default_config(precision=precision, max_iter=4000, **cfg)
    c.item_blocks = item_blocks
    return PS.BatchSolver(c).set_meshes(meshes)


def test_offsets(batch_meshes):
    bs = make_batch(batch_meshes)
    assert list(np.diff(bs.offsets)) == [m.n for m in batch_meshes]
    assert bs.offsets[0] == 0


@pytest.mark.parametrize("norm", [PS.NORM_L2, PS.NORM_INF])
def test_fp64_batch_matches_single(batch_meshes, batch_rhs, norm):
    bs = make_batch(batch_meshes, norm=norm)
    xb, stb = bs.solve(np.concatenate(batch_rhs), rel_tol=1e-12)
    for k, (m, b) in enumerate(zip(batch_meshes, batch_rhs)):
        s = PS.Solver(max_iter=4000, norm=norm).set_mesh(m)
        xs, sts = s.solve(b, rel_tol=1e-12)
        xk = bs.split(xb)[k]
        print(k, m.n_blocks, stb[k], sts)
        assert stb[k]["converged"] and sts["converged"]
        # identical until rounding differences move a breakdown restart
        assert stb[k]["iterations"] <= 1.3 * sts["iterations"] + 10
        assert np.linalg.norm(xk - xs) <= 1e-9 * np.linalg.norm(xs)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
@pytest.mark.parametrize("norm", [PS.NORM_L2, PS.NORM_INF])
def test_batch_reaches_residual(oracle, batch_meshes, batch_rhs, oracle_solutions, precision, norm):
    bs = make_batch(batch_meshes, precision=precision, norm=norm)
    xb, stb = bs.solve(np.concatenate(batch_rhs), rel_tol=1e-10)
    for k, (m, b) in enumerate(zip(batch_meshes, batch_rhs)):
        xk = bs.split(xb)[k]
        tr = oracle.true_residual(oracle.assemble(m), b, xk)
        print(k, m.n_blocks, stb[k], tr["rel2"], tr["relinf"])
        assert stb[k]["converged"], stb[k]
        if precision == PS.MIXED:
            assert stb[k]["refinements"] >= 1
        assert (tr["rel2"] if norm == PS.NORM_L2 else tr["relinf"]) <= 1.05e-10, tr
        xr = oracle_solutions[k]
        assert np.linalg.norm(xk - xr) <= 1e-6 * np.linalg.norm(xr)


def test_mixed_batch_matches_oracle_1e8(batch_meshes, batch_rhs, oracle_solutions):
    bs = make_batch(batch_meshes, precision=PS.MIXED, norm=PS.NORM_L2)
    xb, stb = bs.solve(np.concatenate(batch_rhs), rel_tol=1e-12)
    for k in range(len(batch_meshes)):
        xk, xr = bs.split(xb)[k], oracle_solutions[k]
        assert stb[k]["converged"], stb[k]
        assert np.linalg.norm(xk - xr) <= 1e-8 * np.linalg.norm(xr)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_mean_constraint(oracle, batch_meshes, batch_rhs, precision):
    bs = make_batch(batch_meshes, precision=precision, mean_constraint=1, norm=PS.NORM_L2)
    xb, stb = bs.solve(np.concatenate(batch_rhs), rel_tol=1e-10)
    for k, (m, b) in enumerate(zip(batch_meshes, batch_rhs)):
        s = PS.Solver(max_iter=4000, mean_constraint=1, norm=PS.NORM_L2).set_mesh(m)
        xs, sts = s.solve(b, rel_tol=1e-12)
        xk = bs.split(xb)[k]
        # residual of the system the solver stops on: the mean row replaced
        # by sum h^2 x = 0
        A1 = oracle.assemble(m, mean_constraint=1)
        bm = b.copy()
        bm[A1.mean_row] = 0.0
        tr = oracle.true_residual(A1, bm, xk, mean_mode=1, h2v=oracle.h2(m))
        assert abs(np.dot(np.repeat(oracle.h2(m), 64), xk)) <= 1e-12 * np.abs(xk).max() * m.n
        assert stb[k]["converged"], stb[k]
        assert tr["rel2"] <= 1.05e-10, tr
        assert np.linalg.norm(xk - xs) <= 1e-6 * np.linalg.norm(xs)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_zero_instance(batch_meshes, batch_rhs, precision):
    rhs = list(batch_rhs)
    rhs[1] = np.zeros_like(rhs[1])
    bs = make_batch(batch_meshes, precision=precision, norm=PS.NORM_L2)
    xb, stb = bs.solve(np.concatenate(rhs), rel_tol=1e-10)
    assert stb[1]["converged"] and stb[1]["iterations"] == 0
    assert np.all(bs.split(xb)[1] == 0.0)
    for k in (0, 2, 3, 4):
        assert stb[k]["converged"] and stb[k]["iterations"] > 0


def test_batch_iteration_cap(batch_meshes, batch_rhs):
    bs = PS.BatchSolver(PS.default_config(max_iter=5)).set_meshes(batch_meshes)
    _, stb = bs.solve(np.concatenate(batch_rhs), rel_tol=1e-14)
    for st in stb:
        assert st["iterations"] <= 5 and not st["converged"]


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_deterministic_graph_and_device(batch_meshes, batch_rhs, precision):
    b = np.concatenate(batch_rhs)
    bs = make_batch(batch_meshes, precision=precision)
    x1, s1 = bs.solve(b, rel_tol=1e-10)
    x2, s2 = bs.solve(b, rel_tol=1e-10)
    assert np.array_equal(x1, x2)
    assert [s["iterations"] for s in s1] == [s["iterations"] for s in s2]
    ng = make_batch(batch_meshes, precision=precision, use_graph=0)
    x3, _ = ng.solve(b, rel_tol=1e-10)
    assert np.array_equal(x1, x3)
    db = bs.device_array(b)
    dx = bs.device_array(np.zeros_like(b))
    bs.solve_device(db, dx, rel_tol=1e-10)
    assert np.array_equal(x1, dx.numpy())


def test_batch_item_sizes_and_autotune(batch_meshes, batch_rhs, oracle_solutions):
    b = np.concatenate(batch_rhs)
    ref = None
    for ib in (4, 32):
        bs = make_batch(batch_meshes, precision=PS.MIXED, item_blocks=ib, norm=PS.NORM_L2)
        assert bs.item_blocks() == ib
        x, st = bs.solve(b, rel_tol=1e-10)
        assert all(s["converged"] for s in st)
        if ref is None:
            ref = x
        else:
            assert np.linalg.norm(x - ref) <= 1e-6 * np.linalg.norm(ref)
    bs = make_batch(batch_meshes, precision=PS.MIXED, norm=PS.NORM_L2)
    ib, ms = bs.autotune(bs.device_array(b))
    assert ib in (4, 8, 16, 32, 64) and ms > 0
    assert bs.item_blocks() == ib and bs.xcd_run() in (1, 4, 8, 16, 32)
    x, st = bs.solve(b, rel_tol=1e-10)
    assert all(s["converged"] for s in st)
    for k, xr in enumerate(oracle_solutions):
        assert np.linalg.norm(bs.split(x)[k] - xr) <= 1e-6 * np.linalg.norm(xr)
    # a new handle of the same shape starts from the tuned schedule
    again = make_batch(batch_meshes, precision=PS.MIXED, norm=PS.NORM_L2)
    assert again.item_blocks() == ib and again.xcd_run() == bs.xcd_run()


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_xcd_run_order_is_bitwise_neutral(batch_meshes, batch_rhs, precision):
    """The operator kernels' item order (runs of 1 = plain, 4, 8, 32 items per
    XCD) moves work between workgroups only: every partial sum is stored per
    item and reduced in item order, so the solutions are bit-identical."""
    b = np.concatenate(batch_rhs)
    ref = None
    for run in (1, 4, 8, 32):
        bs = make_batch(batch_meshes, precision=precision, item_blocks=4, norm=PS.NORM_L2, xcd_run=run)
        assert bs.xcd_run() == run
        x, st = bs.solve(b, rel_tol=1e-10)
        assert all(s["converged"] for s in st)
        if ref is None:
            ref = x
        else:
            assert np.array_equal(x, ref), run
    with pytest.raises(PS._lib.SolverError):
        make_batch(batch_meshes, xcd_run=3)


def test_batch_rejects_bad_input(batch_meshes):
    with pytest.raises(PS._lib.SolverError):
        PS.BatchSolver(PS.default_config(max_iter=0))
    bs = make_batch(batch_meshes[:2])
    with pytest.raises(PS._lib.SolverError):
        bs.solve(np.zeros(bs.n), rel_tol=-1.0)


@pytest.mark.parametrize("mc", [0, 1])
def test_batch_spmv_matches_oracle(oracle, batch_meshes, mc):
    bs = make_batch(batch_meshes, mean_constraint=mc)
    rng = np.random.default_rng(12)
    x = rng.normal(size=bs.n)
    y = bs.spmv(x)
    sz = bs.sizes()
    assert sz["n"] == bs.n
    nnz = 0
    for k, m in enumerate(batch_meshes):
        A = oracle.assemble(m, mc)
        nnz += A.nnz
        ref = oracle.spmv(A, bs.split(x)[k], mean_mode=mc, h2v=oracle.h2(m))
        assert np.abs(bs.split(y)[k] - ref).max() <= 1e-14 * np.abs(ref).max()
    assert sz["nnz"] == nnz


def test_batch_copy_time_copies(batch_meshes):
    """The streaming copy used as the bandwidth ceiling moves every unknown."""
    bs = make_batch(batch_meshes)
    x = np.random.default_rng(3).normal(size=bs.n)
    d_x = bs.device_array(x)
    d_y = bs.device_array(np.zeros(bs.n))
    ms = bs.copy_time(d_x, d_y, reps=3)
    assert ms > 0.0
    assert np.array_equal(d_y.numpy(), x)
    with pytest.raises(PS._lib.SolverError):
        bs.copy_time(d_x, d_y, reps=0)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_batch_kernel_time_and_stats(oracle, batch_meshes, batch_rhs, precision):
    """Each batched iteration kernel can be timed alone; the handle solves
    correctly afterwards, and every instance's stats carry its FP64 true
    residual (as the oracle recomputes it) and, with cfg.profile, the phase
    times of the batch."""
    bs = make_batch(batch_meshes, precision=precision, norm=PS.NORM_L2)
    for k in range(8):
        ms = bs.kernel_time(k, reps=3)
        assert 0.0 < ms < 100.0, (k, ms)
    with pytest.raises(PS._lib.SolverError):
        bs.kernel_time(8)
    b = np.concatenate(batch_rhs)
    x, st = bs.solve(b, rel_tol=1e-10)
    for k, (m, bk) in enumerate(zip(batch_meshes, batch_rhs)):
        tr = oracle.true_residual(oracle.assemble(m), bk, bs.split(x)[k])
        assert st[k]["converged"] and tr["rel2"] <= 1.05e-10
        assert abs(st[k]["true_rel2"] - tr["rel2"]) <= 1e-3 * tr["rel2"] + 1e-16
        assert abs(st[k]["true_relinf"] - tr["relinf"]) <= 1e-3 * tr["relinf"] + 1e-16
    bp = make_batch(batch_meshes, precision=precision, norm=PS.NORM_L2, profile=1)
    xp, stp = bp.solve(b, rel_tol=1e-10)
    assert np.array_equal(xp, x)
    ph = stp[0]["ms_phase"]
    assert ph["spmv"] > 0 and ph["prec"] > 0 and ph["scalar"] > 0
    assert 0.5 * stp[0]["ms"] <= sum(ph.values()) <= 1.5 * stp[0]["ms"] + 1.0, (ph, stp[0]["ms"])
