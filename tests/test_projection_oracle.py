"""CPU checks that pin the oracle's restatement of the operators either side
of the pressure solve (oracle/projection.py): closed forms of linear fields,
the ghost rules cell by cell, periodic meshes against numpy.roll, and at
coarse-fine faces the face-form Laplacian against the assembled operator,
exact differences of linear fields, the discrete divergence theorem and
momentum conservation of the correction."""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import workloads as W
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


def _tiny_cf():
    return M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                        (1, 0, 3), (1, 1, 3)])


def _periodic_graded():
    return M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04), (1.0, 0.5, 0.1)],
                    [0.3, 0.15, 0.08], True, True)


CF_MESHES = {"tiny_cf": _tiny_cf, "graded": W.graded_disk_mesh, "periodic": _periodic_graded}


@pytest.mark.parametrize("which", sorted(CF_MESHES))
def test_face_form_laplacian_is_the_assembled_operator(oracle, which):
    """The coarse-fine ghosts of the restatement, put into the face-form
    Laplacian, give the assembled operator's rows (pinned by the known-answer
    row and conservation) at every explicit cell."""
    m = CF_MESHES[which]()
    x = np.random.default_rng(2).normal(size=m.n)
    ax = oracle.assemble(m).to_scipy() @ x
    lf = PR.laplacian_face_form(m, x)
    assert len(lf) == len(PR.explicit_cells(m)) > 0
    err = max(abs(ax[i] - v) for i, v in lf.items())
    assert err <= 1e-14 * np.abs(ax).max()


def _away_from_walls(m, x, y):
    h = m.cell_h()
    Lx = m.extent * m.bpdx / max(m.bpdx, m.bpdy)
    Ly = m.extent * m.bpdy / max(m.bpdx, m.bpdy)
    return (x > h) & (x < Lx - h) & (y > h) & (y < Ly - h)


def test_linear_fields_coarse_fine():
    """Linear velocity and pressure are differentiated exactly at the
    explicit cells too (the ghost interpolation is exact for linear fields
    and the flux correction preserves it), away from the walls."""
    m = W.graded_disk_mesh()
    x, y = m.cell_centers()
    a, b, dt = 0.7, -1.3, 0.01
    z = np.zeros(m.n)
    h = m.cell_h()
    rhs = PR.pressure_rhs(m, a * x + 0.2, b * y - 0.1, z, z, z, dt)
    ex = np.array([PR._gi(*c) for c in PR.explicit_cells(m)])
    keep = ex[_away_from_walls(m, x, y)[ex]]
    assert len(keep) > 100
    ref = h * h * (a + b) / dt
    assert np.allclose(rhs[keep], ref[keep], rtol=1e-11, atol=0)
    al, be = 2.5, -0.4
    uo, vo = PR.pressure_correct(m, al * x + be * y, z, z, dt)
    assert np.allclose(uo[keep], -dt * al, rtol=1e-11, atol=0)
    assert np.allclose(vo[keep], -dt * be, rtol=1e-11, atol=0)


@pytest.mark.parametrize("which", ["graded", "periodic"])
def test_divergence_theorem(which):
    """sum of the RHS over the domain = the flux through its boundary: zero
    for walls (the mirrored normal velocity cancels the face value) and for
    periodic faces; with chi = 0 and with constant chi.  Coarse and fine
    fluxes across every coarse-fine face must be equal and opposite for it
    to hold."""
    m = CF_MESHES[which]()
    rng = np.random.default_rng(3)
    u, v, ud, vd = (rng.normal(size=m.n) for _ in range(4))
    for c in (0.0, 0.3):
        r = PR.pressure_rhs(m, u, v, ud, vd, np.full(m.n, c), 0.01)
        assert abs(r.sum()) <= 1e-15 * np.abs(r).sum() * 10


def test_correction_conserves_momentum_periodic():
    """On a periodic mesh the correction changes sum h^2 u by nothing: its
    face fluxes telescope, coarse-fine faces included."""
    m = _periodic_graded()
    rng = np.random.default_rng(4)
    p = rng.normal(size=m.n)
    z = np.zeros(m.n)
    uo, vo = PR.pressure_correct(m, p, z, z, 0.02)
    w = m.cell_h() ** 2
    assert abs(np.dot(w, uo)) <= 1e-14 * np.dot(w, np.abs(uo))
    assert abs(np.dot(w, vo)) <= 1e-14 * np.dot(w, np.abs(vo))


def test_coarse_fine_plain_cells_unchanged():
    """Cells off the coarse-fine faces are computed exactly as on a mesh
    without such faces: the central formula of the reference."""
    m = _tiny_cf()
    rng = np.random.default_rng(5)
    u, v, ud, vd, chi = (rng.normal(size=m.n) for _ in range(5))
    dt, h = 0.1, m.cell_h()
    rhs = PR.pressure_rhs(m, u, v, ud, vd, chi, dt)
    # block 0 (level 0, covers [0, .5]^2), cell (3, 3): four in-block neighbours
    i = 3 * 8 + 3
    f = 0.5 * h[i] / dt
    ref = f * ((u[i + 1] - u[i - 1]) + (v[i + 8] - v[i - 8]))
    ref = ref + (-f) * chi[i] * ((ud[i + 1] - ud[i - 1]) + (vd[i + 8] - vd[i - 8]))
    assert rhs[i] == ref
