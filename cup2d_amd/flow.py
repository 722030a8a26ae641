"""The projection time step of the reference's pipeline for flow past a fixed
rigid disk, every operator on the GPU (device-resident fields):

    dt = min(0.25 h^2 / (nu + 0.25 h umax), CFL h / umax)  (Simulation.cpp:426-460)
    advection-diffusion, two stages                        (advDiff.cpp:238-282)
    penalization toward the body velocity                  (PressureSingle.cpp:268-322)
    rhs = div u (no deformation velocity)                  (PressureSingle.cpp:324-397)
    p = p_old + dp - <dp>, A dp = rhs - L p_old             (PressureSingle.cpp:827-879)
    u -= dt grad p                                         (PressureSingle.cpp:122-210)

The body is held fixed in the computational frame (the reference's
bFixed = 1): it moves at (us, vs) = -uinf in the laboratory frame, the
fluid's laboratory velocity is what the fields hold, and advection uses
u + uinf (advDiff.cpp:94-97).  With adapt_every > 0 the mesh adapts as the
reference's AdaptTheMesh does (AdaptTheMesh.cpp:73-77: every step of the
first ten, then every adapt_every steps): the per-block vorticity score on
the GPU, the refine / compress decision on the host (mesh.adapt), the
fields moved to the new mesh and the operator re-assembled on the GPU.
"""

from __future__ import annotations

import math

import numpy as np

from . import mesh as M
from . import poisson as PS
from . import workloads as W


class CylinderFlow:
    """Flow past the disk of radius `radius` at `center`, free stream `uinf`,
    viscosity `nu`, penalization `lam` (the Python driver's default 1e6,
    cubismup2d/simulation.py:73), CFL number `cfl`; the pressure solved to
    `rel_tol` relative residual by the handle's solver (FP64 or mixed).
    Adaptation every `adapt_every` steps (0: a static mesh) with the
    refinement / compression tolerances `rtol`, `ctol` on the vorticity of
    the disk launcher (launch/launchDiskAMR.sh:51-52)."""

    def __init__(self, mesh, radius=W.CYL_RADIUS, center=W.CYL_CENTER, uinf=(W.CYL_U, 0.0),
                 nu=W.CYL_NU, lam=1e6, cfl=W.CYL_CFL, precision=PS.MIXED, rel_tol=1e-10,
                 max_iter=4000, device=0, adapt_every=0, rtol=5.0, ctol=0.01):
        self.mesh = mesh
        self.center, self.radius = (float(center[0]), float(center[1])), float(radius)
        self.adapt_every, self.rtol, self.ctol = int(adapt_every), float(rtol), float(ctol)
        self.adaptations = 0
        self.uinf = (float(uinf[0]), float(uinf[1]))
        self.us, self.vs = -self.uinf[0], -self.uinf[1]
        self.nu, self.lam, self.cfl, self.rel_tol = float(nu), float(lam), float(cfl), float(rel_tol)
        self.solver = PS.Solver(device=device, precision=precision, norm=PS.NORM_L2, max_iter=max_iter)
        self.solver.set_mesh(mesh, on_device=True)
        n = mesh.n
        z = np.zeros(n)
        self.u = PS.DeviceArray.from_host(self.solver, z)
        self.v = PS.DeviceArray.from_host(self.solver, z)
        self.p = PS.DeviceArray.from_host(self.solver, z)
        self.rhs = PS.DeviceArray.from_host(self.solver, z)
        self.zero = PS.DeviceArray.from_host(self.solver, z)  # no deformation velocity
        self.chi = PS.DeviceArray.from_host(self.solver, z)
        self.solver.chi_disk_device(center[0], center[1], radius, self.chi)
        self.h_min = float(mesh.h().min())
        self.time = 0.0
        self.steps = 0
        self.last = {}
        self.last_tags = None

    def adapt(self, tags=None):
        """One mesh adaptation: tags from the GPU score (refine above rtol,
        compress below ctol) unless given; returns the tags used."""
        s = self.solver
        if tags is None:
            score = s.block_score_device(self.u, self.v, self.chi, self.rtol)
            tags = np.where(score > self.rtol, 1, np.where(score < self.ctol, -1, 0)).astype(np.int8)
        m2, kind, src = M.adapt(self.mesh, tags)
        if (m2.n_blocks == self.mesh.n_blocks and np.array_equal(m2.level, self.mesh.level)
                and np.array_equal(m2.bx, self.mesh.bx) and np.array_equal(m2.by, self.mesh.by)):
            return tags
        new = [PS.DeviceArray(s, m2.n) for _ in range(3)]
        s.transfer_device(self.mesh, m2, kind, src, [self.u, self.v, self.p], new)
        s.set_mesh(m2, on_device=True)
        self.u, self.v, self.p = new
        self.mesh = m2
        z = np.zeros(m2.n)
        self.rhs = PS.DeviceArray.from_host(s, z)
        self.zero = PS.DeviceArray.from_host(s, z)
        self.chi = PS.DeviceArray.from_host(s, z)
        s.chi_disk_device(self.center[0], self.center[1], self.radius, self.chi)
        self.h_min = float(m2.h().min())
        self.adaptations += 1
        return tags

    def due(self):
        """AdaptTheMesh's schedule (AdaptTheMesh.cpp:73-77)."""
        return self.adapt_every > 0 and (self.steps <= 10 or self.steps % self.adapt_every == 0)

    def time_step(self):
        """calcMaxTimestep (Simulation.cpp:426-460) without ramp-up."""
        umax = self.solver.max_speed_device(self.u, self.v, self.uinf)
        h = self.h_min
        dt_diff = 0.25 * h * h / (self.nu + 0.25 * h * umax)
        dt_adv = h / (umax + 1e-8)
        return min(dt_diff, self.cfl * dt_adv)

    def step(self, dt=None):
        s = self.solver
        self.last_tags = self.adapt() if self.due() else None
        dt = self.time_step() if dt is None else float(dt)
        if not (dt > 0 and math.isfinite(dt)):
            raise RuntimeError(f"time step {dt}")
        s.advect_diffuse_device(self.u, self.v, self.uinf, self.nu, dt)
        s.penalize_device(self.chi, self.us, self.vs, self.lam, dt, self.u, self.v)
        s.pressure_rhs_device(self.u, self.v, self.zero, self.zero, self.chi, dt, self.rhs)
        st = s.pressure_step_device(self.rhs, self.p, rel_tol=self.rel_tol)
        s.pressure_correct_device(self.p, dt, self.u, self.v)
        self.time += dt
        self.steps += 1
        self.last = {"dt": dt, **st}
        return self.last

    def fields(self):
        """(u, v, p) as host arrays."""
        return self.u.numpy(), self.v.numpy(), self.p.numpy()
