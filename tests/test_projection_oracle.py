"""CPU checks that pin the oracle's restatement of the operators either side
of the pressure solve (oracle/projection.py): closed forms of linear fields,
the ghost rules cell by cell, and periodic meshes against numpy.roll."""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from oracle import projection as PR


def _global(m, f):
    """Field of a uniform level-0 mesh as a (ny*8, nx*8) array."""
    g = np.zeros((m.bpdy * 8, m.bpdx * 8))
    blk = np.asarray(f).reshape(-1, 8, 8)
    for b in range(m.n_blocks):
        g[m.by[b] * 8:(m.by[b] + 1) * 8, m.bx[b] * 8:(m.bx[b] + 1) * 8] = blk[b]
    return g


def test_linear_fields_interior():
    m = M.uniform(4, 3, 1.0)
    x, y = m.cell_centers()
    a, b, dt = 0.7, -1.3, 0.01
    u, v = a * x + 0.2, b * y - 0.1
    z = np.zeros(m.n)
    rhs = _global(m, PR.pressure_rhs(m, u, v, z, z, z, dt))
    h = m.h()[0]
    inner = rhs[1:-1, 1:-1]  # cells away from the domain faces
    assert np.allclose(inner, h * h * (a + b) / dt, rtol=1e-12, atol=0)
    # p = alpha x + beta y: the correction is -dt (alpha, beta) away from the faces
    al, be = 2.5, -0.4
    uo, vo = PR.pressure_correct(m, al * x + be * y, z, z, dt)
    assert np.allclose(_global(m, uo)[1:-1, 1:-1], -dt * al, rtol=1e-12)
    assert np.allclose(_global(m, vo)[1:-1, 1:-1], -dt * be, rtol=1e-12)


def test_ghost_rules_one_block():
    m = M.uniform(1, 1, 1.0)
    rng = np.random.default_rng(0)
    u, v, p = (rng.normal(size=64) for _ in range(3))
    z = np.zeros(64)
    dt = 0.5
    h = m.h()[0]
    U, V, P = u.reshape(8, 8), v.reshape(8, 8), p.reshape(8, 8)
    rhs = PR.pressure_rhs(m, u, v, z, z, z, dt).reshape(8, 8)
    f = 0.5 * h / dt
    # west face: ghost u = -u(0); east face: ghost u = -u(7); same for v in y
    assert rhs[3, 0] == f * ((U[3, 1] + U[3, 0]) + (V[4, 0] - V[2, 0]))
    assert rhs[3, 7] == f * ((-U[3, 7] - U[3, 6]) + (V[4, 7] - V[2, 7]))
    assert rhs[0, 3] == f * ((U[0, 4] - U[0, 2]) + (V[1, 3] + V[0, 3]))
    assert rhs[7, 3] == f * ((U[7, 4] - U[7, 2]) + (-V[7, 3] - V[6, 3]))
    uo, vo = PR.pressure_correct(m, p, z, z, dt)
    uo, vo = uo.reshape(8, 8), vo.reshape(8, 8)
    pf, ih2 = -0.5 * dt * h, 1.0 / h / h
    # zero-gradient pressure ghosts
    assert uo[3, 0] == (pf * (P[3, 1] - P[3, 0])) * ih2
    assert uo[3, 7] == (pf * (P[3, 7] - P[3, 6])) * ih2
    assert vo[0, 3] == (pf * (P[1, 3] - P[0, 3])) * ih2
    assert vo[7, 3] == (pf * (P[7, 3] - P[6, 3])) * ih2


def test_periodic_against_roll():
    m = M.uniform(4, 2, 2.0, True, True)
    rng = np.random.default_rng(1)
    u, v, ud, vd, chi, p = (rng.normal(size=m.n) for _ in range(6))
    dt = 0.03
    h = m.h()[0]
    G = lambda f: _global(m, f)
    f = 0.5 * h / dt
    dx = lambda g: np.roll(g, -1, axis=1) - np.roll(g, 1, axis=1)
    dy = lambda g: np.roll(g, -1, axis=0) - np.roll(g, 1, axis=0)
    ref = f * (dx(G(u)) + dy(G(v))) + (-f) * G(chi) * (dx(G(ud)) + dy(G(vd)))
    assert np.array_equal(G(PR.pressure_rhs(m, u, v, ud, vd, chi, dt)), ref)
    uo, vo = PR.pressure_correct(m, p, u, v, dt)
    pf, ih2 = -0.5 * dt * h, 1.0 / h / h
    assert np.array_equal(G(uo), G(u) + pf * dx(G(p)) * ih2)
    assert np.array_equal(G(vo), G(v) + pf * dy(G(p)) * ih2)


def test_coarse_fine_mesh_rejected():
    m = M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                     (1, 0, 3), (1, 1, 3)])
    z = np.zeros(m.n)
    with pytest.raises(ValueError):
        PR.pressure_rhs(m, z, z, z, z, z, 1.0)
