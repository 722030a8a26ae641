"""GPU parity tests: the HIP solver (through the C ABI) against the CPU
oracle on the same seeded inputs.

Tolerances: the operator and preconditioner are FP64 with a different
summation order than the oracle, so single applications agree to a few ulps
(1e-14 relative to the largest entry).  Solves are compared after the same
h^2-weighted mean removal: relative L2 difference <= 1e-8 against the oracle
solved to FP64 stagnation, and true relative residual <= 1e-10.
"""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _tiny_cf():
    return M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                        (1, 0, 3), (1, 1, 3)])


MESHES = {
    "tiny_cf": _tiny_cf,
    "uniform16x8": W.uniform_16x8_mesh,
    "graded": W.graded_disk_mesh,
    "amr6_small": lambda: W.four_disk_amr6_mesh(target=3000),
    "periodic": lambda: M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)],
                                 [0.3, 0.15, 0.08], True, True),
}


@pytest.fixture(scope="module")
def solvers():
    cache = {}

    def get(name, mc=0, **cfg):
        key = (name, mc, tuple(sorted(cfg.items())))
        if key not in cache:
            m = MESHES[name]()
            s = PS.Solver(mean_constraint=mc, **cfg)
            s.set_mesh(m)
            cache[key] = (m, s)
        return cache[key]
    return get


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("mc", [0, 1])
def test_spmv_matches_oracle(oracle, solvers, name, mc):
    m, s = solvers(name, mc)
    A = oracle.assemble(m, mc)
    h2 = oracle.h2(m)
    rng = np.random.default_rng(11)
    x = rng.normal(size=m.n)
    ref = oracle.spmv(A, x, mean_mode=mc, h2v=h2)
    scale = np.abs(ref).max()
    for kind in (0, 1):
        y = s.spmv(x, kind)
        assert np.abs(y - ref).max() <= 1e-14 * scale, (name, kind)


@pytest.mark.parametrize("name", ["tiny_cf", "graded"])
def test_precond_matches_oracle(oracle, solvers, name):
    m, s = solvers(name)
    P = oracle.pinv()
    r = np.random.default_rng(3).normal(size=m.n)
    ref = (P @ r.reshape(-1, 64).T).T.reshape(-1)
    z = s.precond(r)
    assert np.abs(z - ref).max() <= 1e-14 * np.abs(ref).max()


def _solve_case(oracle, solvers, name, mc, norm, seed=1, rel_tol=1e-10):
    m, s = solvers(name, mc, norm=norm, max_iter=4000)
    A = oracle.assemble(m, mc)
    h2 = oracle.h2(m)
    b = W.vortex_rhs(m, seed)
    bb = b.copy()
    if mc:
        bb[A.mean_row] = 0.0
    x, st = s.solve(b, None, tol=0.0, rel_tol=rel_tol, max_restarts=100)
    # oracle solved to stagnation
    xr, sr = oracle.bicgstab(A, bb, mean_mode=mc, h2v=h2, rel_tol=1e-13, norm="l2", max_iter=4000)
    xr = oracle.mean_remove(h2, xr)
    return m, A, h2, bb, x, st, xr, sr


@pytest.mark.parametrize("name", ["tiny_cf", ("uniform16x8"), "graded", "amr6_small", "periodic"])
@pytest.mark.parametrize("norm", [PS.NORM_INF, PS.NORM_L2])
def test_solve_reaches_residual(oracle, solvers, name, norm):
    """Stopping at 1e-10 relative residual: the independently recomputed true
    residual meets it, the result is h^2-mean-free, and it is within the
    conditioning-limited distance of the oracle's solution."""
    m, A, h2, bb, x, st, xr, sr = _solve_case(oracle, solvers, name, 0, norm)
    print(name, norm, st)
    assert st["converged"], st
    tr = oracle.true_residual(A, bb, x)
    assert tr["rel2"] <= 1.05e-10 if norm == PS.NORM_L2 else tr["relinf"] <= 1.05e-10, tr
    assert abs(np.dot(np.repeat(h2, 64), x)) <= 1e-12 * np.abs(x).max() * h2.sum()
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert err <= 1e-6, err


@pytest.mark.parametrize("name", ["tiny_cf", ("uniform16x8"), "graded", "amr6_small", "periodic"])
def test_solve_matches_oracle_1e8(oracle, solvers, name):
    """Solved to 1e-12 relative residual, the pressure agrees with the FP64
    oracle to 1e-8 relative L2 after the same mean removal."""
    m, A, h2, bb, x, st, xr, sr = _solve_case(oracle, solvers, name, 0, PS.NORM_L2, rel_tol=1e-12)
    print(name, st, sr)
    tr = oracle.true_residual(A, bb, x)
    assert tr["rel2"] <= 2e-12, tr
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert err <= 1e-8, err


@pytest.mark.parametrize("name", ["tiny_cf", "graded"])
def test_solve_mean_constraint_matches_oracle(oracle, solvers, name):
    m, A, h2, bb, x, st, xr, sr = _solve_case(oracle, solvers, name, 1, PS.NORM_L2, rel_tol=1e-12)
    assert st["converged"], st
    tr = oracle.true_residual(A, bb, x, mean_mode=0)  # mean-free x: mean row residual is 0 too
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    assert err <= 1e-8, err


def test_iteration_trajectory_close_to_oracle(oracle, solvers):
    """Same algorithm, same inputs: iteration and restart counts of the GPU
    solve track the oracle's reference-algorithm run."""
    m, s = solvers(("uniform16x8"), 0, norm=PS.NORM_INF)
    A = oracle.assemble(m)
    b = W.vortex_rhs(m, 1)
    _, st = s.solve(b, rel_tol=1e-8, max_restarts=100)
    _, sr = oracle.bicgstab(A, b, rel_tol=1e-8, norm="inf")
    assert abs(st["iterations"] - sr["iterations"]) <= max(3, 0.1 * sr["iterations"]), (st, sr)
    assert abs(st["err_init"] - sr["err_init"]) <= 1e-15 * sr["err_init"] + 1e-300


def test_reference_eps_stagnates_like_oracle(oracle):
    """eps = 1e-21 (BiCGSTAB.cu:427) stalls on the h^2-scaled RHS in both."""
    m = W.uniform_16x8_mesh()
    s = PS.Solver(eps=1e-21, norm=PS.NORM_L2)
    s.set_mesh(m)
    b = W.vortex_rhs(m, 1)
    _, st = s.solve(b, rel_tol=1e-10, max_restarts=100)
    A = oracle.assemble(m)
    _, sr = oracle.bicgstab(A, b, rel_tol=1e-10, norm="l2", eps=1e-21)
    assert not st["converged"] and not sr["converged"]
    assert st["iterations"] == 1000 or st["restarts"] >= 1


def test_zero_rhs_returns_mean_free_guess(solvers):
    m, s = solvers(("uniform16x8"))
    x0 = np.random.default_rng(2).normal(size=m.n)
    b = np.zeros(m.n)
    # x0 is not a solution of A x = 0 unless constant: use a constant guess
    x0[:] = 3.0
    x, st = s.solve(b, x0, rel_tol=1e-10)
    assert st["iterations"] == 0
    assert np.abs(x).max() < 1e-14


def test_iteration_cap_returns_best_iterate(oracle, solvers):
    m, s0 = solvers("graded")
    s = PS.Solver(max_iter=7)
    s.set_mesh(m)
    b = W.vortex_rhs(m, 4)
    x, st = s.solve(b, rel_tol=1e-12)
    assert st["iterations"] == 7 and not st["converged"]
    A = oracle.assemble(m)
    tr = oracle.true_residual(A, b, x)
    assert tr["relinf"] <= 1.0001 * st["err_final"] / st["err_init"] + 1e-12


def test_deterministic_and_graph_equivalent(solvers):
    m, s = solvers("graded")
    b = W.vortex_rhs(m, 7)
    x1, st1 = s.solve(b, rel_tol=1e-10)
    x2, st2 = s.solve(b, rel_tol=1e-10)
    assert np.array_equal(x1, x2) and st1["iterations"] == st2["iterations"]
    s3 = PS.Solver(use_graph=0)
    s3.set_mesh(m)
    x3, st3 = s3.solve(b, rel_tol=1e-10)
    assert np.array_equal(x1, x3)
    s4 = PS.Solver(chunk=5)
    s4.set_mesh(m)
    x4, _ = s4.solve(b, rel_tol=1e-10)
    assert np.array_equal(x1, x4)


def test_device_solve_equals_host_solve(solvers):
    m, s = solvers("graded")
    b = W.vortex_rhs(m, 8)
    xh, sth = s.solve(b, rel_tol=1e-10)
    db = PS.DeviceArray.from_host(s, b)
    dx = PS.DeviceArray.from_host(s, np.zeros(m.n))
    std = s.solve_device(db, dx, rel_tol=1e-10)
    assert np.array_equal(dx.numpy(), xh)
    res = s.residual_device(db, dx)
    assert res["relinf"] <= 1.5e-10


def test_poisson_solver_interface(oracle):
    """makePoissonSolver / solve(input, output) as PressureSingle calls it
    (PressureSingle.cpp:836-851): x0 = 0, output receives the mean-free Δp."""
    m = W.graded_disk_mesh()
    sim = PS.SimulationData(poissonSolver="cuda_iterative", PoissonTol=0.0, PoissonTolRel=1e-10, step=20)
    sim.tmp = PS.ScalarGrid(m)
    sim.pres = PS.ScalarGrid(m)
    b = W.vortex_rhs(m, 9)
    sim.tmp.data[...] = b.reshape(-1, 8, 8)
    solver = PS.makePoissonSolver(sim)
    solver.solve(sim.tmp, sim.pres)
    assert solver.last_stats["converged"]
    assert not sim.pres.UpdateFluxCorrection
    A = oracle.assemble(m)
    tr = oracle.true_residual(A, b, sim.pres.data.reshape(-1))
    assert tr["relinf"] <= 1.5e-10
    # input untouched
    assert np.array_equal(sim.tmp.data.reshape(-1), b)
    # the operator is reused while the mesh is unchanged (ExpAMRSolver.cpp:370-380)
    assert solver.assemblies == 1
    sim.pres.data[...] = 0.0  # x0 = 0 for every solve (PressureSingle.cpp:845)
    solver.solve(sim.tmp, sim.pres)
    assert solver.assemblies == 1 and solver.last_stats["converged"]
    # a new mesh object (the same blocks) and the mesh-changed flag re-assemble
    m2 = W.graded_disk_mesh()
    sim.pres = PS.ScalarGrid(m2, UpdateFluxCorrection=False)
    solver.solve(sim.tmp, sim.pres)
    assert solver.assemblies == 2
    sim.pres.UpdateFluxCorrection = True
    sim.pres.data[...] = 0.0
    solver.solve(sim.tmp, sim.pres)
    assert solver.assemblies == 3 and not sim.pres.UpdateFluxCorrection
    tr = oracle.true_residual(A, b, sim.pres.data.reshape(-1))
    assert tr["relinf"] <= 1.5e-10


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_kernel_time_then_solve(oracle, precision):
    """Timing the iteration kernels alone leaves the handle usable: the next
    solve starts afresh and is correct."""
    m = W.uniform_16x8_mesh()
    s = PS.Solver(precision=precision, norm=PS.NORM_L2, max_iter=4000).set_mesh(m)
    for k in range(5):
        ms = s.kernel_time(k, reps=5)
        assert 0.0 < ms < 100.0
    b = W.vortex_rhs(m, 3)
    x, st = s.solve(b, rel_tol=1e-10)
    assert st["converged"]
    tr = oracle.true_residual(oracle.assemble(m), b, x)
    assert tr["rel2"] <= 1.05e-10
    with pytest.raises(PS._lib.SolverError):
        s.kernel_time(5)


def test_unfused_comparator_matches_oracle(oracle):
    """schedule=1: the one-operation-per-launch FP64 loop of the reference
    (BiCGSTAB.cu:412-620) reaches the same solution as the oracle; in the
    infinity norm with the reference's own stopping test."""
    for name in ("graded", "tiny_cf"):
        m = MESHES[name]()
        A = oracle.assemble(m)
        b = W.vortex_rhs(m, 17)
        xr, _ = oracle.bicgstab(A, b, rel_tol=1e-13, norm="l2", max_iter=6000)
        xr = oracle.mean_remove(oracle.h2(m), xr)
        for norm in (PS.NORM_L2, PS.NORM_INF):
            s = PS.Solver(schedule=1, norm=norm, max_iter=4000).set_mesh(m)
            # the stopping test reads the recurrence residual, as the
            # reference's; at 1e-10 the true residual meets the tolerance
            x, st = s.solve(b, rel_tol=1e-10)
            tr = oracle.true_residual(A, b, x)
            assert st["converged"], (name, st)
            assert (tr["rel2"] if norm == PS.NORM_L2 else tr["relinf"]) <= 1.05e-10, tr
            # solved to 1e-12 (where the recurrence and the true residual
            # part by rounding: within 1e-11), the solution is the oracle's
            x, st = s.solve(b, rel_tol=1e-12)
            tr = oracle.true_residual(A, b, x)
            assert st["converged"], (name, st)
            assert (tr["rel2"] if norm == PS.NORM_L2 else tr["relinf"]) <= 1e-11, tr
            assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
            # the stats' own true residual agrees with the oracle's
            assert abs(st["true_rel2"] - tr["rel2"]) <= 1e-3 * tr["rel2"] + 1e-16
    # the mean-row form
    m = MESHES["graded"]()
    s1 = PS.Solver(schedule=1, mean_constraint=1, norm=PS.NORM_L2, max_iter=4000).set_mesh(m)
    b = W.vortex_rhs(m, 18)
    x, st = s1.solve(b, rel_tol=1e-10)
    A1 = oracle.assemble(m, 1)
    bm = b.copy()
    bm[A1.mean_row] = 0.0
    tr = oracle.true_residual(A1, bm, x, mean_mode=1, h2v=oracle.h2(m))
    assert st["converged"] and tr["rel2"] <= 1.05e-10
    with pytest.raises(PS._lib.SolverError):
        PS.Solver(schedule=1, precision=PS.MIXED)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_profile_phases_and_true_residual(oracle, precision):
    """cfg.profile: per-phase device times that add up to the solve, and the
    same solution as the graph-replayed solve; every solve reports its FP64
    true residual."""
    m = MESHES["graded"]()
    A = oracle.assemble(m)
    b = W.vortex_rhs(m, 19)
    cfg = dict(precision=precision, norm=PS.NORM_L2, max_iter=4000)
    x0, st0 = PS.Solver(**cfg).set_mesh(m).solve(b, rel_tol=1e-10)
    x1, st1 = PS.Solver(profile=1, **cfg).set_mesh(m).solve(b, rel_tol=1e-10)
    assert np.array_equal(x0, x1)
    assert all(v == 0.0 for v in st0["ms_phase"].values())
    ph = st1["ms_phase"]
    assert ph["spmv"] > 0 and ph["prec"] > 0 and ph["update"] > 0 and ph["setup"] > 0
    assert 0.5 * st1["ms"] <= sum(ph.values()) <= 1.5 * st1["ms"] + 1.0, (ph, st1["ms"])
    tr = oracle.true_residual(A, b, x0)
    for st in (st0, st1):
        assert abs(st["true_rel2"] - tr["rel2"]) <= 1e-3 * tr["rel2"] + 1e-16
        assert abs(st["true_relinf"] - tr["relinf"]) <= 1e-3 * tr["relinf"] + 1e-16
    # the unfused comparator's phases
    if precision == PS.FP64:
        _, st2 = PS.Solver(profile=1, schedule=1, **cfg).set_mesh(m).solve(b, rel_tol=1e-10)
        assert st2["converged"] and st2["ms_phase"]["spmv"] > 0 and st2["ms_phase"]["scalar"] > 0
