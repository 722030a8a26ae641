"""CPU restatement of the two operators either side of the pressure solve,
for meshes without coarse-fine faces.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
(cup2d_amd) never imports it.

* pressure_rhs follows updatePressureRHS (source/Operators/PressureSingle.cpp:
  324-349): rhs = facDiv ((u_E - u_W) + (v_N - v_S)) - facDiv chi ((ud_E -
  ud_W) + (vd_N - vd_S)), facDiv = 0.5 h / dt, in that order of operations.
* pressure_correct follows pressureCorrectionKernel and
  PressureSingle::pressureCorrection (PressureSingle.cpp:122-210): tmp =
  pFac (p_E - p_W), pFac = -0.5 dt h; u += tmp * ih2, ih2 = 1 / h / h.
* Ghosts at non-periodic domain faces (the labs of source/Definitions.h):
  the velocity lab mirrors the boundary cell with the face-normal component
  negated, for walls and free-slip faces alike (Definitions.h:84-117; the
  tangential component, which is where the two differ, is not read by
  either operator); the pressure lab is zero-gradient (ghost = boundary
  cell), as the explicit matrix's wall rows (ExpAMRSolver.cpp:301) assume.
  That lab's code is Cubism's BlockLabNeumann (Definitions.h:199-222 only
  calls it; the submodule is absent), so the pressure ghost at
  non-periodic domain faces is parity unpinned against the reference:
  it rests on the matrix's convention.

Pinned in tests/test_projection_oracle.py by closed forms: linear fields
(exact central differences), the ghost rules on a 1-block mesh checked cell
by cell, and periodic meshes against numpy.roll.
"""

from __future__ import annotations

import numpy as np

BS = 8


def _h(mesh) -> np.ndarray:
    h0 = mesh.extent / (float(max(mesh.bpdx, mesh.bpdy)) * BS)
    return np.array([h0 / float(1 << int(l)) for l in mesh.level], dtype=np.float64)


def _neighbours(mesh):
    """Same-level neighbour block of each face (Xm, Xp, Ym, Yp), -1 at a
    domain face; raises on a coarse-fine face."""
    where = {(int(l), int(i), int(j)): b for b, (l, i, j) in
             enumerate(zip(mesh.level, mesh.bx, mesh.by))}
    nb = np.full((mesh.n_blocks, 4), -1, dtype=np.int64)
    for b, (l, i, j) in enumerate(zip(mesh.level, mesh.bx, mesh.by)):
        l, i, j = int(l), int(i), int(j)
        nx, ny = mesh.bpdx << l, mesh.bpdy << l
        for f, (di, dj) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            I, J = i + di, j + dj
            if (not 0 <= I < nx and not mesh.periodic_x) or (not 0 <= J < ny and not mesh.periodic_y):
                continue
            I, J = I % nx, J % ny
            if (l, I, J) not in where:
                raise ValueError("coarse-fine face: not covered by this restatement")
            nb[b, f] = where[(l, I, J)]
    return nb


def _lab(mesh, f, xs, ys):
    """(n_blocks, 10, 10) array of f with a one-cell halo: neighbours' cells,
    or xs / ys times the boundary cell at domain faces (index [b, y+1, x+1])."""
    nb = _neighbours(mesh)
    blk = np.asarray(f, dtype=np.float64).reshape(-1, BS, BS)
    lab = np.zeros((blk.shape[0], BS + 2, BS + 2))
    lab[:, 1:-1, 1:-1] = blk
    for b in range(blk.shape[0]):
        xm, xp, ym, yp = nb[b]
        lab[b, 1:-1, 0] = blk[xm, :, BS - 1] if xm >= 0 else xs * blk[b, :, 0]
        lab[b, 1:-1, -1] = blk[xp, :, 0] if xp >= 0 else xs * blk[b, :, BS - 1]
        lab[b, 0, 1:-1] = blk[ym, BS - 1, :] if ym >= 0 else ys * blk[b, 0, :]
        lab[b, -1, 1:-1] = blk[yp, 0, :] if yp >= 0 else ys * blk[b, BS - 1, :]
    return lab


def _diffs(lab):
    """E - W and N - S of every cell."""
    return lab[:, 1:-1, 2:] - lab[:, 1:-1, :-2], lab[:, 2:, 1:-1] - lab[:, :-2, 1:-1]


def pressure_rhs(mesh, u, v, udef_u, udef_v, chi, dt) -> np.ndarray:
    h = _h(mesh)[:, None, None]
    facDiv = 0.5 * h / dt
    du, _ = _diffs(_lab(mesh, u, -1.0, 1.0))
    _, dv = _diffs(_lab(mesh, v, 1.0, -1.0))
    dud, _ = _diffs(_lab(mesh, udef_u, -1.0, 1.0))
    _, dvd = _diffs(_lab(mesh, udef_v, 1.0, -1.0))
    c = np.asarray(chi, dtype=np.float64).reshape(-1, BS, BS)
    t = facDiv * (du + dv)
    t = t + (-facDiv) * c * (dud + dvd)
    return t.reshape(-1)


def pressure_correct(mesh, p, u, v, dt):
    h = _h(mesh)[:, None, None]
    pFac = -0.5 * dt * h
    ih2 = 1.0 / h / h
    dpx, dpy = _diffs(_lab(mesh, p, 1.0, 1.0))
    tu = pFac * dpx
    tv = pFac * dpy
    uo = np.asarray(u, dtype=np.float64).reshape(-1, BS, BS) + tu * ih2
    vo = np.asarray(v, dtype=np.float64).reshape(-1, BS, BS) + tv * ih2
    return uo.reshape(-1), vo.reshape(-1)
