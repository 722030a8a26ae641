"""GPU tests of the flow operators of the projection time step (advection-
diffusion, penalization, the disk's characteristic function, the largest
velocity) against the oracle's restatement (oracle/flow.py), bit for bit
(the HIP kernels compute the same IEEE operations in the same order), and a
short run of the whole time step past a cylinder against the same steps on
the CPU (oracle flow operators, projection operators and FP64 BiCGSTAB)."""

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import flow as FLOW
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W
from oracle import flow as FL
from oracle import projection as PR

pytestmark = pytest.mark.gpu

MESHES = {
    "uniform_walls": lambda: M.uniform(4, 3, 1.0),
    "periodic": lambda: M.uniform(3, 2, 1.0, True, True),
    "tiny_cf": lambda: M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2),
                                                   (1, 1, 2), (1, 0, 3), (1, 1, 3)]),
    "cylinder3": lambda: W.cylinder_mesh(levels=3, target=300),
    "graded_periodic": lambda: M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04),
                                                       (1.0, 0.5, 0.1)], [0.3, 0.15, 0.08], True, True),
}


def _fields(m, seed):
    rng = np.random.default_rng(seed)
    x, y = m.cell_centers()
    u = 0.2 * np.sin(3 * x) * np.cos(2 * y) + 0.05 * rng.normal(size=m.n)
    v = -0.1 * np.cos(2 * x) * np.sin(3 * y) + 0.05 * rng.normal(size=m.n)
    return u, v


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("on_device", [False, True])
def test_advect_diffuse_matches_oracle(name, on_device):
    m = MESHES[name]()
    u, v = _fields(m, len(name))
    s = PS.Solver().set_mesh(m, on_device=on_device)
    for uinf, nu, dt in (((0.2, 0.0), 2e-5, 1e-3), ((-0.1, 0.3), 1e-2, 5e-4)):
        uo, vo = s.advect_diffuse(u, v, uinf, nu, dt)
        ru, rv = FL.advect_diffuse(m, u, v, uinf, nu, dt)
        assert np.array_equal(uo, ru) and np.array_equal(vo, rv), (name, uinf)


def test_penalize_chi_and_max_speed_match_oracle():
    m = W.cylinder_mesh(levels=3, target=300)
    s = PS.Solver().set_mesh(m)
    d_chi = PS.DeviceArray(s, m.n)
    s.chi_disk_device(0.3, 0.5, 0.05, d_chi)
    chi = d_chi.numpy()
    rchi = FL.chi_disk(m, 0.3, 0.5, 0.05)
    assert np.allclose(chi, rchi, rtol=0, atol=4e-16), np.abs(chi - rchi).max()
    assert (chi > 0).sum() > 100
    u, v = _fields(m, 3)
    d_u, d_v = PS.DeviceArray.from_host(s, u), PS.DeviceArray.from_host(s, v)
    d_rc = PS.DeviceArray.from_host(s, rchi)
    s.penalize_device(d_rc, -0.2, 0.05, 1e6, 1e-3, d_u, d_v)
    ru, rv = FL.penalize(u, v, rchi, -0.2, 0.05, 1e6, 1e-3)
    assert np.array_equal(d_u.numpy(), ru) and np.array_equal(d_v.numpy(), rv)
    assert s.max_speed_device(d_u, d_v, (0.2, 0.0)) == FL.max_speed(ru, rv, (0.2, 0.0))
    # argument checks
    with pytest.raises(_lib.SolverError):
        s.penalize_device(d_rc, 0.0, 0.0, 1e6, 0.0, d_u, d_v)
    with pytest.raises(_lib.SolverError):
        s.advect_diffuse_device(d_u, d_u, (0.0, 0.0), 1e-3, 1e-3)


def _oracle_step(oracle, m, A, h2, u, v, p, chi, flow, dt):
    """The same time step on the CPU: oracle flow and projection operators,
    the pressure by the oracle's FP64 BiCGSTAB (to 1e-12) with the mean
    removal of the pressure step."""
    u, v = FL.advect_diffuse(m, u, v, flow.uinf, flow.nu, dt)
    u, v = FL.penalize(u, v, chi, flow.us, flow.vs, flow.lam, dt)
    z = np.zeros(m.n)
    rhs = PR.pressure_rhs(m, u, v, z, z, chi, dt)
    b = rhs - oracle.spmv(A, p)
    dp, st = oracle.bicgstab(A, b, rel_tol=1e-12, norm="l2", max_iter=6000)
    w = np.repeat(h2, 64)
    avg = np.dot(w, dp) / np.sum(w)
    p = dp + (p - avg)
    u, v = PR.pressure_correct(m, p, u, v, dt)
    return u, v, p


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_cylinder_steps_match_cpu_run(oracle, precision):
    """Five time steps of flow past the cylinder (r = 0.05, U = 0.2, Re =
    1000, CFL 0.4, 3-level static mesh): velocity and pressure agree with the
    CPU run to 1e-7 relative (the pressure solves differ by their tolerance)."""
    m = W.cylinder_mesh(levels=3, target=300)
    flow = FLOW.CylinderFlow(m, precision=precision, rel_tol=1e-12, max_iter=6000)
    A = oracle.assemble(m)
    h2 = oracle.h2(m)
    chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
    u = np.zeros(m.n)
    v = np.zeros(m.n)
    p = np.zeros(m.n)
    for k in range(5):
        dt = FL.time_step(m, u, v, flow.uinf, flow.nu, flow.cfl)
        st = flow.step(dt)
        assert st["converged"], (k, st)
        u, v, p = _oracle_step(oracle, m, A, h2, u, v, p, chi, flow, dt)
    gu, gv, gp = flow.fields()
    for g, r in ((gu, u), (gv, v), (gp, p)):
        assert np.linalg.norm(g - r) <= 1e-7 * np.linalg.norm(r)
    # the body moves with the frame: the velocity inside is the body's
    inside = chi > 0.99
    assert np.abs(gu[inside] + flow.uinf[0]).max() < 0.05
