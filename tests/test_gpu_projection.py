"""GPU tests of the operators either side of the pressure solve
(cup2d_ps_pressure_rhs_device, cup2d_ps_pressure_correct_device) against the
oracle's restatement (oracle/projection.py), on uniform meshes and on
graded meshes with coarse-fine faces (face form with the flux correction
there).  Both compute the same IEEE operations in the same order (the HIP
kernels without contraction), so the results are compared bit for bit."""

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W
from oracle import projection as PR

pytestmark = pytest.mark.gpu

MESHES = {
    "uniform16x8_walls": lambda: M.uniform(16, 8, 2.0),
    "periodic_xy": lambda: M.uniform(4, 2, 2.0, True, True),
    "periodic_x": lambda: M.uniform(3, 5, 1.0, True, False),
    "one_block": lambda: M.uniform(1, 1, 1.0),
    "level2_uniform": lambda: M.from_blocks(1, 1, 3, 1.0, [(2, i, j) for j in range(4) for i in range(4)]),
    "tiny_cf": lambda: M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2),
                                                   (1, 1, 2), (1, 0, 3), (1, 1, 3)]),
    "graded3": W.graded_disk_mesh,
    "graded_periodic": lambda: M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04),
                                                       (1.0, 0.5, 0.1)], [0.3, 0.15, 0.08], True, True),
}


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("on_device", [False, True])
def test_projection_operators_match_oracle(name, on_device):
    m = MESHES[name]()
    rng = np.random.default_rng(len(name))
    u, v, ud, vd, p = (rng.normal(size=m.n) for _ in range(5))
    chi = rng.uniform(0, 1, size=m.n)
    dt = 0.0125
    s = PS.Solver().set_mesh(m, on_device=on_device)
    rhs = s.pressure_rhs(u, v, ud, vd, chi, dt)
    assert np.array_equal(rhs, PR.pressure_rhs(m, u, v, ud, vd, chi, dt)), name
    uo, vo = s.pressure_correct(p, u, v, dt)
    ru, rv = PR.pressure_correct(m, p, u, v, dt)
    assert np.array_equal(uo, ru) and np.array_equal(vo, rv), name


def test_projection_operators_reject_bad_dt_and_aliases():
    u = M.uniform(2, 2, 1.0)
    s = PS.Solver().set_mesh(u)
    z = np.zeros(u.n)
    with pytest.raises(_lib.SolverError):
        s.pressure_rhs(z, z, z, z, z, 0.0)
    with pytest.raises(_lib.SolverError):
        s.pressure_correct(z, z, z, -1.0)
    # rhs written in place over an input is rejected (single and batch)
    d = [PS.DeviceArray.from_host(s, z) for _ in range(5)]
    for k in range(5):
        with pytest.raises(_lib.SolverError, match="alias"):
            s.pressure_rhs_device(*d, 1.0, d[k])
    bs = PS.BatchSolver().set_meshes([u, u])
    db = [bs.device_array(np.zeros(bs.n)) for _ in range(5)]
    lib = _lib.load()
    for k in range(5):
        assert lib.cup2d_psb_pressure_rhs_device(bs.h, *[a.ptr for a in db], 1.0, db[k].ptr) == 1


@pytest.mark.parametrize("on_device", [False, True])
def test_coarse_fine_divergence_theorem_and_linear_fields(on_device):
    """On the graded mesh: the RHS sums to zero over the walled domain
    (coarse and fine fluxes equal and opposite), and linear fields are
    differentiated exactly at the coarse-fine cells away from the walls."""
    m = W.graded_disk_mesh()
    s = PS.Solver().set_mesh(m, on_device=on_device)
    rng = np.random.default_rng(8)
    u, v, ud, vd = (rng.normal(size=m.n) for _ in range(4))
    z = np.zeros(m.n)
    r = s.pressure_rhs(u, v, ud, vd, z, 0.01)
    assert abs(r.sum()) <= 1e-14 * np.abs(r).sum()
    x, y = m.cell_centers()
    h = m.cell_h()
    a, b, dt = 0.7, -1.3, 0.01
    r = s.pressure_rhs(a * x + 0.2, b * y - 0.1, z, z, z, dt)
    ex = np.array([PR._gi(*c) for c in PR.explicit_cells(m)])
    keep = ex[(x[ex] > h[ex]) & (x[ex] < 2 - h[ex]) & (y[ex] > h[ex]) & (y[ex] < 1 - h[ex])]
    assert np.allclose(r[keep], (h * h * (a + b) / dt)[keep], rtol=1e-11, atol=0)
    uo, vo = s.pressure_correct(2.5 * x - 0.4 * y, z, z, dt)
    assert np.allclose(uo[keep], -dt * 2.5, rtol=1e-11) and np.allclose(vo[keep], dt * 0.4, rtol=1e-11)


def test_projection_round_trip_reduces_divergence():
    """rhs of u*, the pressure step, the correction: the corrected velocity's
    divergence RHS is smaller than u*'s (the collocated projection is not
    exact, but it removes most of the divergence of a smooth field)."""
    m = M.uniform(8, 8, 1.0, True, True)
    x, y = m.cell_centers()
    u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.3 * np.cos(2 * np.pi * x)
    v = np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y) + 0.2 * np.sin(4 * np.pi * y)
    z = np.zeros(m.n)
    dt = 0.01
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000).set_mesh(m)
    rhs0 = s.pressure_rhs(u, v, z, z, z, dt)
    rhs0 -= rhs0.mean()  # periodic: compatible
    p, st = s.pressure_step(rhs0, z, rel_tol=1e-12)
    uo, vo = s.pressure_correct(p, u, v, dt)
    rhs1 = s.pressure_rhs(uo, vo, z, z, z, dt)
    assert st["converged"]
    assert np.linalg.norm(rhs1) < 0.05 * np.linalg.norm(rhs0)


@pytest.mark.parametrize("on_device", [False, True])
def test_batch_projection_operators_match_oracle(on_device):
    """All instances in one call: each instance's slice equals the oracle's."""
    meshes = [MESHES[n]() for n in ("uniform16x8_walls", "one_block", "periodic_xy", "periodic_x",
                                    "level2_uniform", "tiny_cf", "graded3", "graded_periodic")]
    bs = PS.BatchSolver().set_meshes(meshes, on_device=on_device)
    rng = np.random.default_rng(7)
    u, v, ud, vd, p = (rng.normal(size=bs.n) for _ in range(5))
    chi = rng.uniform(0, 1, size=bs.n)
    dt = 0.02
    rhs = bs.pressure_rhs(u, v, ud, vd, chi, dt)
    uo, vo = bs.pressure_correct(p, u, v, dt)
    for k, m in enumerate(meshes):
        sl = slice(bs.offsets[k], bs.offsets[k + 1])
        assert np.array_equal(rhs[sl], PR.pressure_rhs(m, u[sl], v[sl], ud[sl], vd[sl], chi[sl], dt)), k
        ru, rv = PR.pressure_correct(m, p[sl], u[sl], v[sl], dt)
        assert np.array_equal(uo[sl], ru) and np.array_equal(vo[sl], rv), k
