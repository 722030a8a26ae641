"""Synthetic Poisson workloads: the meshes and right-hand sides the solver is
measured and tested on.

Right-hand side: the discrete divergence of a seeded velocity field made of
16 Gaussian (Lamb-Oseen type) vortices, taken the way the reference builds
the pressure RHS (central differences scaled by h/2 with dt = 1,
source/Operators/PressureSingle.cpp:336-349) but with the analytic field at
the neighbour centres, then projected onto the range of the Neumann operator
(1^T b = 0) so that the singular system is consistent.
"""

from __future__ import annotations

import math

import numpy as np

from . import mesh as meshmod

SIGMA = 0.05
N_VORTICES = 16


def vortex_velocity(x, y, centers, gammas, sigma=SIGMA):
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for (cx, cy), g in zip(centers, gammas):
        dx = x - cx
        dy = y - cy
        r2 = dx * dx + dy * dy
        r2s = np.where(r2 > 1e-300, r2, 1.0)
        f = np.where(r2 > 1e-300, g / (2 * math.pi * r2s) * (1 - np.exp(-r2 / (2 * sigma * sigma))),
                     g / (4 * math.pi * sigma * sigma))
        u += -dy * f
        v += dx * f
    return u, v


def vortex_rhs(m: meshmod.Mesh, seed: int, n_vortices=N_VORTICES, sigma=SIGMA, sources=None):
    """b_i = (h_i/2)[(u(x+h)-u(x-h)) + (v(y+h)-v(y-h))] (+ h_i^2 * sources),
    then b <- b - (sum b / sum h^2) h^2."""
    rng = np.random.default_rng(seed)
    Lx, Ly = m.domain()
    centers = np.stack([rng.uniform(0, Lx, n_vortices), rng.uniform(0, Ly, n_vortices)], axis=1)
    gammas = rng.normal(0.0, 1.0, n_vortices)
    x, y = m.cell_centers()
    h = m.cell_h()
    up, _ = vortex_velocity(x + h, y, centers, gammas, sigma)
    um, _ = vortex_velocity(x - h, y, centers, gammas, sigma)
    _, vp = vortex_velocity(x, y + h, centers, gammas, sigma)
    _, vm = vortex_velocity(x, y - h, centers, gammas, sigma)
    b = 0.5 * h * ((up - um) + (vp - vm))
    if sources is not None:
        for (sx, sy, a, ss) in sources:
            b += h * h * a * np.exp(-((x - sx) ** 2 + (y - sy) ** 2) / (2 * ss * ss))
    h2 = h * h
    b -= (b.sum() / h2.sum()) * h2
    return b


def disk_sources(seed, cx, cy, r, count=8, sigma=0.02):
    rng = np.random.default_rng(seed + 7919)
    ang = rng.uniform(0, 2 * math.pi, count)
    amp = rng.normal(0.0, 1.0, count)
    return [(cx + r * math.cos(t), cy + r * math.sin(t), a, sigma) for t, a in zip(ang, amp)]


# ---------------------------------------------------------------------------
# Meshes of the measured configurations.
# ---------------------------------------------------------------------------

def uniform_16x8_mesh():
    """Uniform 16x8 blocks of 8x8 cells, domain [0,2]x[0,1] (128x64 cells)."""
    return meshmod.uniform(16, 8, extent=2.0)


# The stated rule of the 3-level case (level 2 within 0.15, level 1 within 0.3
# of the disk centre) yields 320 blocks, far from its stated ~1.5k blocks /
# ~100k cells; the radii are therefore scaled by 5 (0.75 and 1.5 from the
# centre), which gives 1448 blocks / 92.7k cells.
GRADED_RADIUS_SCALE = 5.0


def graded_disk_mesh(scale=GRADED_RADIUS_SCALE):
    cx, cy, r = 0.5, 0.5, 0.1
    bands = [0.3 * scale - r, 0.15 * scale - r]
    return meshmod.graded(16, 8, 3, 2.0, [(cx, cy, r)], bands)


def _bisect_mesh(make, target, tol=0.02, lo=0.05, hi=8.0, iters=40):
    """Find a band scale whose mesh has N_b within tol of target."""
    best = None
    for _ in range(iters):
        s = math.sqrt(lo * hi)
        m = make(s)
        err = (m.n_blocks - target) / target
        if best is None or abs(err) < abs(best[1]):
            best = (m, err, s)
        if abs(err) <= tol:
            break
        if m.n_blocks < target:
            lo = s
        else:
            hi = s
    m, err, s = best
    m.meta["band_scale"] = s
    m.meta["target_blocks"] = target
    return m


def multi_disk_mesh(seed, ndisks, rmin, rmax, levels, target, tol=0.02):
    rng = np.random.default_rng(seed)
    disks = [(float(rng.uniform(0.15, 1.85)), float(rng.uniform(0.15, 0.85)),
              float(rng.uniform(rmin, rmax))) for _ in range(ndisks)]

    def make(s):
        bands = [s * 0.4 * 0.5 ** lv for lv in range(levels - 1)]
        return meshmod.graded(16, 8, levels, 2.0, disks, bands)

    return _bisect_mesh(make, target, tol)


def four_disk_amr6_mesh(seed=3, target=16384):
    """6 levels refined around 4 seeded disks (r in [0.02,0.08]), ~16k blocks."""
    return multi_disk_mesh(seed, 4, 0.02, 0.08, 6, target)


def batch64_meshes(seed=4, count=64, nmin=512, nmax=8192, tol=0.05):
    """`count` independent meshes, N_b log-uniform in [nmin, nmax] (seeded),
    each refined around its own random disk to within tol of its N_b."""
    rng = np.random.default_rng(seed)
    targets = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), count))).astype(int)
    out = []
    for k, t in enumerate(targets):
        out.append(multi_disk_mesh(seed * 1000 + k, 1, 0.05, 0.15, 5, int(t), tol))
    return out


def workload_rhs(m, seed, with_sources=False):
    src = None
    if with_sources:
        src = disk_sources(seed, 0.5, 0.5, 0.1)
    return vortex_rhs(m, seed, sources=src)


# Flow past a cylinder (the projection time-step workload): disk r = 0.05 at
# (0.3, 0.5) in the [0, 2] x [0, 1] domain (16 x 8 level-0 blocks), U = 0.2,
# Re = U 2r / nu = 1000, CFL 0.4.
CYL_CENTER = (0.3, 0.5)
CYL_RADIUS = 0.05
CYL_U = 0.2
CYL_RE = 1000.0
CYL_NU = CYL_U * 2 * CYL_RADIUS / CYL_RE
CYL_CFL = 0.4


def cylinder_mesh(levels=5, target=8192, tol=0.02):
    """A static mesh refined in bands around the cylinder (levels up to
    `levels`), ~`target` blocks."""
    cx, cy = CYL_CENTER

    def make(s):
        bands = [s * 0.6 * 0.5 ** lv for lv in range(levels - 1)]
        return meshmod.graded(16, 8, levels, 2.0, [(cx, cy, CYL_RADIUS)], bands)

    return _bisect_mesh(make, target, tol)
