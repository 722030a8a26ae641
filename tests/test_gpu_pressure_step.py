"""GPU tests of the device-resident pressure step (cup2d_ps_pressure_step_device,
PressureSingle.cpp:842-879): b = rhs - L p_old, dp solved from 0, p = dp + (p_old - <dp>_h2).

Checked against the oracle's pieces on the same inputs (its plain operator,
its BiCGSTAB solved to FP64 stagnation, its h^2 mean removal): relative L2
difference of the pressure increment <= 1e-8.  Invariants that need no
oracle: L p = rhs to the solve's tolerance (L p_old cancels), the increment
is h^2-mean-free, and p_old = 0 reduces the step to a plain solve of rhs.
"""

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _tiny_cf():
    return M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                        (1, 0, 3), (1, 1, 3)])


MESHES = {
    "tiny_cf": _tiny_cf,
    "graded": W.graded_disk_mesh,
    "amr6_small": lambda: W.four_disk_amr6_mesh(target=3000),
    "periodic": lambda: M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)],
                                 [0.3, 0.15, 0.08], True, True),
}


def _p_old(m, seed):
    """A smooth old pressure plus noise, of the increment's magnitude."""
    x, y = m.cell_centers()
    rng = np.random.default_rng(seed)
    return 1e-4 * (np.cos(3 * x + 1) * np.sin(2 * y) + 0.1 * rng.normal(size=m.n))


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("mc", [0, 1])
@pytest.mark.parametrize("on_device", [False, True])
def test_pressure_step_matches_oracle(oracle, name, mc, on_device):
    m = MESHES[name]()
    rhs = W.vortex_rhs(m, 5)
    p_old = _p_old(m, 6)
    s = PS.Solver(mean_constraint=mc, norm=PS.NORM_L2, max_iter=4000).set_mesh(m, on_device=on_device)
    p, st = s.pressure_step(rhs, p_old, rel_tol=1e-12)
    assert st["converged"], st
    # oracle: b = rhs - L p_old (plain operator), mean row of b = 0 with the constraint
    h2 = oracle.h2(m)
    L = oracle.assemble(m, 0)
    b = rhs - oracle.spmv(L, p_old)
    A = oracle.assemble(m, mc)
    if mc:
        b[A.mean_row] = 0.0
    dp, sr = oracle.bicgstab(A, b, mean_mode=mc, h2v=h2, rel_tol=1e-13, norm="l2", max_iter=4000)
    dp = oracle.mean_remove(h2, dp)
    inc = p - p_old
    assert np.linalg.norm(inc - dp) <= 1e-8 * np.linalg.norm(dp), name
    # the increment is h^2-mean-free
    w = np.repeat(h2, 64)
    assert abs(np.dot(w, inc)) <= 1e-12 * np.abs(inc).max() * w.sum()
    # L p = rhs up to the solve's tolerance on b (L p_old cancels); with the
    # constraint the mean row's equation is replaced, so it is left out
    res = rhs - oracle.spmv(L, p)
    if mc:
        res[A.mean_row] = 0.0
    assert np.linalg.norm(res) <= 2e-11 * np.linalg.norm(b) + 1e-14 * np.linalg.norm(rhs), name


def test_pressure_step_from_zero_is_a_solve():
    m = W.graded_disk_mesh()
    rhs = W.vortex_rhs(m, 7)
    s = PS.Solver(norm=PS.NORM_L2).set_mesh(m)
    p, st = s.pressure_step(rhs, np.zeros(m.n), rel_tol=1e-10)
    x, st2 = s.solve(rhs, rel_tol=1e-10)
    assert st["iterations"] == st2["iterations"]
    assert np.abs(p - x).max() <= 1e-14 * np.abs(x).max()


def test_pressure_step_keeps_rhs_and_repeats():
    """rhs is not modified; a second step from the first step's pressure has
    nothing left to do beyond the solve tolerance (its b is at the first
    solve's residual level, so it is stopped by an absolute tolerance: a
    relative one would ask for a reduction below rounding)."""
    m = W.four_disk_amr6_mesh(target=3000)
    rhs = W.vortex_rhs(m, 8)
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000).set_mesh(m)
    d_rhs = PS.DeviceArray.from_host(s, rhs)
    d_p = PS.DeviceArray.from_host(s, _p_old(m, 9))
    st1 = s.pressure_step_device(d_rhs, d_p, rel_tol=1e-11)
    p1 = d_p.numpy()
    st2 = s.pressure_step_device(d_rhs, d_p, tol=1e-10 * st1["err_init"], rel_tol=0.0)
    p2 = d_p.numpy()
    assert np.array_equal(d_rhs.numpy(), rhs)
    assert st1["converged"] and st2["converged"]
    # both pressures meet the tolerance: they agree to the pressure bar
    # (1e-8 relative L2; the increment is A^-1 of a residual at 1e-11 of b)
    assert np.linalg.norm(p2 - p1) <= 1e-8 * np.linalg.norm(p1)


def test_pressure_step_errors():
    m = _tiny_cf()
    s = PS.Solver()
    with pytest.raises(_lib.SolverError):  # no mesh yet
        s.pressure_step(np.zeros(m.n), np.zeros(m.n))
    s.set_mesh(m)
    d_a = PS.DeviceArray.from_host(s, np.zeros(m.n))
    with pytest.raises(_lib.SolverError):  # rhs and p alias
        s.pressure_step_device(d_a, d_a)
    with pytest.raises(_lib.SolverError):
        s.pressure_step(np.zeros(m.n), np.zeros(m.n), rel_tol=-1.0)


@pytest.mark.parametrize("mc", [0, 1])
def test_batch_pressure_step_matches_oracle(oracle, mc):
    """The batched pressure step, each instance against the oracle's pieces
    (as the single-mesh test), instances of different sizes in one call.  The
    step solves to 1e-13 here, as the oracle does: two solutions at 1e-12 of
    b differ by up to the condition number times that, which on the 6-level
    instance is the whole 1e-8 bar (the oracle's own OpenMP reductions move
    its solution by ~1e-9 from run to run)."""
    meshes = [MESHES[n]() for n in ("graded", "tiny_cf", "amr6_small", "periodic")]
    bs = PS.BatchSolver(mean_constraint=mc, norm=PS.NORM_L2, max_iter=4000).set_meshes(meshes, on_device=True)
    rhs = np.concatenate([W.vortex_rhs(m, 20 + k) for k, m in enumerate(meshes)])
    p_old = np.concatenate([_p_old(m, 30 + k) for k, m in enumerate(meshes)])
    d_rhs = bs.device_array(rhs)
    d_p = bs.device_array(p_old)
    st = bs.pressure_step_device(d_rhs, d_p, rel_tol=1e-13)
    p = d_p.numpy()
    assert np.array_equal(d_rhs.numpy(), rhs)
    off = bs.offsets
    for k, m in enumerate(meshes):
        assert st[k]["converged"], (k, st[k])
        sl = slice(off[k], off[k + 1])
        h2 = oracle.h2(m)
        L = oracle.assemble(m, 0)
        b = rhs[sl] - oracle.spmv(L, p_old[sl])
        A = oracle.assemble(m, mc)
        if mc:
            b[A.mean_row] = 0.0
        dp, _ = oracle.bicgstab(A, b, mean_mode=mc, h2v=h2, rel_tol=1e-13, norm="l2", max_iter=4000)
        dp = oracle.mean_remove(h2, dp)
        inc = p[sl] - p_old[sl]
        assert np.linalg.norm(inc - dp) <= 1e-8 * np.linalg.norm(dp), k
        w = np.repeat(h2, 64)
        assert abs(np.dot(w, inc)) <= 1e-12 * np.abs(inc).max() * w.sum(), k


def test_batch_pressure_step_errors():
    bs = PS.BatchSolver()
    with pytest.raises(_lib.SolverError):  # no meshes yet
        bs.pressure_step_device(PS.DeviceArray(bs.mem, 64), PS.DeviceArray(bs.mem, 64))
    bs.set_meshes([_tiny_cf()])
    d_a = bs.device_array(np.zeros(bs.n))
    with pytest.raises(_lib.SolverError):  # rhs and p alias
        bs.pressure_step_device(d_a, d_a)


def test_pressure_step_mixed_precision(oracle):
    """The mixed-precision handle: the same step to the same FP64 tolerance."""
    m = W.graded_disk_mesh()
    rhs = W.vortex_rhs(m, 12)
    p_old = _p_old(m, 13)
    s = PS.Solver(precision=PS.MIXED, norm=PS.NORM_L2, max_iter=6000).set_mesh(m, on_device=True)
    p, st = s.pressure_step(rhs, p_old, rel_tol=1e-10)
    assert st["converged"], st
    L = oracle.assemble(m, 0)
    b = rhs - oracle.spmv(L, p_old)
    res = rhs - oracle.spmv(L, p)
    assert np.linalg.norm(res) <= 1.05e-10 * np.linalg.norm(b)
