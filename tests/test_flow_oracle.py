"""CPU checks that pin the flow operators' restatement (oracle/flow.py):
WENO5 exactness for quadratics and its order on smooth data, the upwind
bias, ghost exactness for linear fields across coarse-fine faces,
conservation of the flux-corrected diffusion, constant fields left
unchanged by the two-stage update, the penalization's closed form and the
disk's characteristic function."""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import workloads as W
from oracle import flow as FL
from oracle import projection as PR


def _periodic_graded():
    return M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04), (1.0, 0.5, 0.1)],
                    [0.3, 0.15, 0.08], True, True)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_weno_exact_for_quadratics(sign):
    """Every candidate stencil reconstructs a quadratic exactly, so any
    convex combination does: fp - fm is q(i + 1/2) - q(i - 1/2) = b + 2 c i
    (unit spacing) whatever the smoothness weights, for either upwind side."""
    rng = np.random.default_rng(0)
    a, b, c = rng.normal(size=3)
    i = np.arange(-5.0, 6.0)
    q = lambda x: a + b * x + c * x * x
    U = np.full(i.shape, sign)
    d = FL.derivative(U, q(i - 3), q(i - 2), q(i - 1), q(i), q(i + 1), q(i + 2), q(i + 3))
    assert np.allclose(d, b + 2 * c * i, rtol=0, atol=1e-12 * (abs(b) + abs(c) * 10))


def test_weno_order_on_smooth_data():
    """On point values of a smooth periodic field (fp - fm) / h is the
    derivative to fifth order (the reconstruction of cell averages, applied
    to point values of sin(kx), returns face values of sin(kx) / sinc, whose
    difference is exactly h k cos(kx)): halving h divides the error by ~32,
    which the linear weights (0.1, 0.6, 0.3) alone give."""
    errs = []
    for n in (32, 64):
        h = 1.0 / n
        x = np.arange(n) * h
        k = 2 * np.pi
        f = lambda s: np.sin(k * s)
        d = FL.derivative(np.ones(n), *(f(x + j * h) for j in (-3, -2, -1, 0, 1, 2, 3))) / h
        errs.append(np.abs(d - k * np.cos(k * x)).max())
    assert errs[0] / errs[1] > 24, errs


def test_weno_upwind_no_new_extrema():
    """A step reconstructed from the upwind side stays within the data's
    range (no overshoot at the face), for both signs of the velocity."""
    u = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0])
    for k in range(1, 5):
        win = np.roll(u, -k)[:5]
        fp = FL.weno5_plus(*win)
        fm = FL.weno5_minus(*win)
        assert -1e-12 <= fp <= 1 + 1e-12 and -1e-12 <= fm <= 1 + 1e-12


@pytest.mark.parametrize("which", ["graded", "periodic"])
def test_ghosts_exact_for_linear_fields(which):
    """Every ghost across a same-level or coarse-fine face of a linear field
    is the field at the ghost's centre (prolongation: cell plus gradient;
    restriction: the average of four fine cells)."""
    m = W.graded_disk_mesh() if which == "graded" else _periodic_graded()
    x, y = m.cell_centers()
    a, b, c = 0.3, -1.7, 2.2
    lab = FL.lab3(m, a + b * x + c * y, 1.0, 1.0)
    h = m.cell_h().reshape(-1, 64)[:, 0]
    fc = PR.faces(m)
    Lx = m.extent * m.bpdx / max(m.bpdx, m.bpdy)
    Ly = m.extent * m.bpdy / max(m.bpdx, m.bpdy)
    x0 = x.reshape(-1, 8, 8)[:, 0, 0] - h / 2
    y0 = y.reshape(-1, 8, 8)[:, 0, 0] - h / 2
    checked = 0
    for bk in range(m.n_blocks):
        for f_, kind in enumerate(fc[bk]):
            if kind[0] not in ("coarse", "fine"):
                continue
            for d in range(1, 4):
                for t in range(8):
                    if f_ == 0:
                        gy, gx = 3 + t, 3 - d
                    elif f_ == 1:
                        gy, gx = 3 + t, 10 + d
                    elif f_ == 2:
                        gy, gx = 3 - d, 3 + t
                    else:
                        gy, gx = 10 + d, 3 + t
                    px = x0[bk] + (gx - 3 + 0.5) * h[bk]
                    py = y0[bk] + (gy - 3 + 0.5) * h[bk]
                    if which == "periodic":
                        continue_ok = 0 <= px <= Lx and 0 <= py <= Ly
                        if not continue_ok:
                            continue
                    exp = a + b * px + c * py
                    assert abs(lab[bk, gy, gx] - exp) <= 1e-12, (bk, f_, d, t, lab[bk, gy, gx], exp)
                    checked += 1
    assert checked > 500


def test_diffusion_conserved_across_coarse_fine_faces():
    """Periodic graded mesh: the flux-corrected diffusive term sums to zero
    (sum h^2 du/dt = sum tmp), for random fields."""
    m = _periodic_graded()
    rng = np.random.default_rng(1)
    u, v = rng.normal(size=m.n), rng.normal(size=m.n)
    tu, tv = FL.rhs_adv_diff(m, u, v, (0.0, 0.0), 1e-3, 0.01, advection=False)
    assert abs(tu.sum()) <= 1e-14 * np.abs(tu).sum()
    assert abs(tv.sum()) <= 1e-14 * np.abs(tv).sum()


def test_constant_field_unchanged():
    """A constant velocity has no advective or diffusive change, across
    coarse-fine faces too (the ghosts of a constant are the constant);
    walls mirror the normal component, so use a periodic mesh."""
    m = _periodic_graded()
    u = np.full(m.n, 0.37)
    v = np.full(m.n, -0.21)
    uo, vo = FL.advect_diffuse(m, u, v, (0.2, 0.05), 2e-5, 1e-3)
    assert np.allclose(uo, u, rtol=0, atol=1e-15) and np.allclose(vo, v, rtol=0, atol=1e-15)


def test_penalize_closed_form():
    rng = np.random.default_rng(2)
    u, v = rng.normal(size=100), rng.normal(size=100)
    chi = np.concatenate([np.zeros(25), np.full(25, 0.3), np.full(25, 0.7), np.ones(25)])
    lam, dt, us, vs = 1e6, 1e-3, -0.2, 0.05
    uo, vo = FL.penalize(u, v, chi, us, vs, lam, dt)
    a = 1.0 / (1.0 + lam * dt)
    assert np.array_equal(uo[:50], u[:50]) and np.array_equal(vo[:50], v[:50])
    assert np.allclose(uo[50:], a * u[50:] + (1 - a) * us, rtol=1e-15)
    assert np.allclose(vo[50:], a * v[50:] + (1 - a) * vs, rtol=1e-15)


def test_chi_disk():
    """chi = 1 inside and 0 outside beyond one cell of the circle, between 0
    and 1 on the band, and its integral is the disk's area to O(h r)."""
    m = M.graded(4, 4, 4, 1.0, [(0.5, 0.5, 0.1)], [0.3, 0.2, 0.15], False, False)
    chi = FL.chi_disk(m, 0.5, 0.5, 0.1)
    x, y = m.cell_centers()
    h = m.cell_h()
    d = 0.1 - np.hypot(x - 0.5, y - 0.5)
    assert np.all(chi[d > h] == 1.0) and np.all(chi[d < -h] == 0.0)
    assert chi.min() >= -1e-12 and chi.max() <= 1 + 1e-12
    area = np.sum(chi * h * h)
    assert abs(area - np.pi * 0.01) <= 2 * np.pi * 0.1 * h.min() * 0.5
