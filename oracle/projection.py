"""CPU restatement of the two operators either side of the pressure solve.

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
* Cells on a coarse-fine block face ("explicit cells") are computed in face
  form: the operator is the sum over the cell's faces Xm, Xp, Ym, Yp of
  s_f * flux_f (s = -1 on the low faces, +1 on the high ones), flux_f being
  the face value the reference's kernels store for the flux correction
  (PressureSingle.cpp:350-397 for the RHS, :139-185 for the correction):
  facDiv (nb + self) - (facDiv chi) (ud_nb + ud_self), or pFac (p_nb +
  p_self).  At a face whose neighbour is one level coarser, nb is the
  coarse-fine ghost of the explicit operator (ExpAMRSolver.cpp:100-128,
  D1/D2 ExpAMRSolver.h:268-300): 2/3 close - 1/5 far + 8/15 (c + sT D1(c) +
  D2(c)).  At a face whose neighbours are one level finer, flux_f is the sum
  of the two fine cells' fluxes across it (each with its own facDiv, chi and
  ghost): the flux correction that makes coarse and fine fluxes equal and
  opposite.  Cubism's own ghost interpolation and FluxCorrection are absent
  here, so equality with them is parity unpinned; what pins this
  restatement is the mathematics: the face-form Laplacian with the same
  ghosts is the assembled operator, linear fields are differentiated
  exactly at coarse-fine cells, and the sum of the RHS over a walled domain
  vanishes (the discrete divergence theorem).

Pinned in tests/test_projection_oracle.py by closed forms: linear fields
(exact central differences), the ghost rules on a 1-block mesh checked cell
by cell, periodic meshes against numpy.roll, and on graded meshes the
checks above.
"""

from __future__ import annotations

import numpy as np

BS = 8


def _h(mesh) -> np.ndarray:
    h0 = mesh.extent / (float(max(mesh.bpdx, mesh.bpdy)) * BS)
    return np.array([h0 / float(1 << int(l)) for l in mesh.level], dtype=np.float64)


def _where(mesh):
    return {(int(l), int(i), int(j)): b for b, (l, i, j) in
            enumerate(zip(mesh.level, mesh.bx, mesh.by))}


def faces(mesh):
    """Per block and face (Xm, Xp, Ym, Yp): ("wall",), ("same", nb),
    ("coarse", cb) or ("fine", f_low, f_high) (ExpAMRSolver.cpp:221-249,
    :141-177)."""
    where = _where(mesh)
    out = []
    for b, (l, i, j) in enumerate(zip(mesh.level, mesh.bx, mesh.by)):
        l, i, j = int(l), int(i), int(j)
        nx, ny = mesh.bpdx << l, mesh.bpdy << l
        fb = []
        for f, (di, dj) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            I, J = i + di, j + dj
            if (not 0 <= I < nx and not mesh.periodic_x) or (not 0 <= J < ny and not mesh.periodic_y):
                fb.append(("wall",))
                continue
            I, J = I % nx, J % ny
            if (l, I, J) in where:
                fb.append(("same", where[(l, I, J)]))
            elif (l - 1, I >> 1, J >> 1) in where:
                fb.append(("coarse", where[(l - 1, I >> 1, J >> 1)]))
            else:
                kids = []
                for hh in range(2):
                    if f < 2:
                        ci, cj = 2 * I + (1 if f == 0 else 0), 2 * J + hh
                    else:
                        ci, cj = 2 * I + hh, 2 * J + (1 if f == 2 else 0)
                    if (l + 1, ci, cj) not in where:
                        raise ValueError(f"block {b} face {f}: mesh not 2:1 balanced")
                    kids.append(where[(l + 1, ci, cj)])
                fb.append(("fine", kids[0], kids[1]))
        out.append(fb)
    return out


def _neighbours(mesh):
    """Same-level neighbour block of each face (Xm, Xp, Ym, Yp), -1 at a
    domain face, -2 at a coarse-fine face."""
    nb = np.full((mesh.n_blocks, 4), -1, dtype=np.int64)
    for b, fb in enumerate(faces(mesh)):
        for f, fc in enumerate(fb):
            nb[b, f] = fc[1] if fc[0] == "same" else (-2 if fc[0] != "wall" else -1)
    return nb


def _gi(b, ix, iy):
    return 64 * b + 8 * iy + ix


def ghost(x, close, far, c, t, step, sT):
    """Coarse-fine ghost value 2/3 close - 1/5 far + 8/15 (c + sT D1(c) +
    D2(c)) (ExpAMRSolver.cpp:100-128; D1/D2 one-sided at the coarse
    tangential indices 3, 7 (backward) and 0, 4 (forward), central
    elsewhere, ExpAMRSolver.h:144-147, :206-209, :268-300), evaluated in
    this order of operations; step(d) is the global index of the coarse
    cell d steps along the face."""
    xc = x[c]
    if t in (3, 7):
        a, bb = x[step(-2)], x[step(-1)]
        d1 = ((1.0 / 8.0) * a + (-1.0 / 2.0) * bb) + (3.0 / 8.0) * xc
        d2 = ((1.0 / 32.0) * a + (-1.0 / 16.0) * bb) + (1.0 / 32.0) * xc
    elif t in (0, 4):
        a, bb = x[step(2)], x[step(1)]
        d1 = ((-1.0 / 8.0) * a + (1.0 / 2.0) * bb) + (-3.0 / 8.0) * xc
        d2 = ((1.0 / 32.0) * a + (-1.0 / 16.0) * bb) + (1.0 / 32.0) * xc
    else:
        a, bb = x[step(-1)], x[step(1)]
        d1 = (-1.0 / 8.0) * a + (1.0 / 8.0) * bb
        d2 = ((1.0 / 32.0) * a + (1.0 / 32.0) * bb) + (-1.0 / 16.0) * xc
    return ((2.0 / 3.0) * x[close] + (-1.0 / 5.0) * x[far]) + (8.0 / 15.0) * ((xc + sT * d1) + d2)


def fine_ghost(mesh, x, b, ix, iy, f, cb):
    """Ghost of fine cell (ix, iy) of block b across face f, whose neighbour
    is the coarser block cb (ExpAMRSolver.cpp:150-160)."""
    xf = f < 2
    if xf:
        ixc = 7 if f == 0 else 0
        iyc = iy // 2 + (4 if int(mesh.by[b]) & 1 else 0)
        inward = _gi(b, ix + 1 if f == 0 else ix - 1, iy)
        step = lambda d: _gi(cb, ixc, iyc + d)
        t = iyc
    else:
        iyc = 7 if f == 2 else 0
        ixc = ix // 2 + (4 if int(mesh.bx[b]) & 1 else 0)
        inward = _gi(b, ix, iy + 1 if f == 2 else iy - 1)
        step = lambda d: _gi(cb, ixc + d, iyc)
        t = ixc
    sT = -1.0 if (iy if xf else ix) % 2 == 0 else 1.0
    return ghost(x, _gi(b, ix, iy), inward, _gi(cb, ixc, iyc), t, step, sT)


def fine_cells(b, ix, iy, f, kids):
    """The two fine cells (close, far, sT) behind face f of coarse cell
    (ix, iy) of block b (ExpAMRSolver.cpp:161-176)."""
    xf = f < 2
    tc = iy if xf else ix
    fbk = kids[1] if tc >= 4 else kids[0]
    t0 = (tc % 4) * 2
    out = []
    for h in range(2):
        t = t0 + h
        if f == 0:
            close, far = _gi(fbk, 7, t), _gi(fbk, 6, t)
        elif f == 1:
            close, far = _gi(fbk, 0, t), _gi(fbk, 1, t)
        elif f == 2:
            close, far = _gi(fbk, t, 7), _gi(fbk, t, 6)
        else:
            close, far = _gi(fbk, t, 0), _gi(fbk, t, 1)
        out.append((close, far, -1.0 if h == 0 else 1.0))
    return out


def coarse_side_ghost(x, b, ix, iy, f, close, far, sT):
    """Ghost of the fine cell `close` facing coarse cell (ix, iy) of block b
    across face f of the coarse block: the same interpolation, with the
    coarse cell's tangential neighbours in its own block."""
    xf = f < 2
    step = (lambda d: _gi(b, ix, iy + d)) if xf else (lambda d: _gi(b, ix + d, iy))
    return ghost(x, close, far, _gi(b, ix, iy), iy if xf else ix, step, sT)


def explicit_cells(mesh):
    """(block, ix, iy) of every cell on a coarse-fine block face, in mesh
    order (block, then cell)."""
    fc = faces(mesh)
    out = []
    for b in range(mesh.n_blocks):
        cf = [fc[b][f][0] in ("coarse", "fine") for f in range(4)]
        for iy in range(8):
            for ix in range(8):
                edge = (ix == 0, ix == 7, iy == 0, iy == 7)
                if any(e and c for e, c in zip(edge, cf)):
                    out.append((b, ix, iy))
    return out


def _face_flux(mesh, fc, b, ix, iy, f, h, coef, comp, dcomp, chi, ghost_sign):
    """flux_f of cell (ix, iy) of block b across face f: coef (nb + self)
    - (coef chi) (dnb + dself) (dcomp None: the first term alone).
    coef(h) is facDiv or pFac of a cell of size h; ghost_sign: the wall
    ghost of comp is ghost_sign * self (-1: mirrored normal velocity, +1:
    zero-gradient pressure)."""
    i = _gi(b, ix, iy)
    edge = (ix == 0, ix == 7, iy == 0, iy == 7)
    di = ((-1, 0), (1, 0), (0, -1), (0, 1))[f]

    def one(cellv, nbv, dself, dnb, c, k):
        t = c * (nbv + cellv)
        if dcomp is not None:
            t = t - (c * chi[k]) * (dnb + dself)
        return t

    hc = h[b]
    if not edge[f]:
        j = _gi(b, ix + di[0], iy + di[1])
        return one(comp[i], comp[j], None if dcomp is None else dcomp[i],
                   None if dcomp is None else dcomp[j], coef(hc), i)
    kind = fc[b][f]
    if kind[0] == "wall":
        return one(comp[i], ghost_sign * comp[i], None if dcomp is None else dcomp[i],
                   None if dcomp is None else ghost_sign * dcomp[i], coef(hc), i)
    if kind[0] == "same":
        s = kind[1]
        jx = {0: 7, 1: 0}.get(f, ix)
        jy = {2: 7, 3: 0}.get(f, iy)
        j = _gi(s, jx, jy)
        return one(comp[i], comp[j], None if dcomp is None else dcomp[i],
                   None if dcomp is None else dcomp[j], coef(hc), i)
    if kind[0] == "coarse":
        g = fine_ghost(mesh, comp, b, ix, iy, f, kind[1])
        gd = None if dcomp is None else fine_ghost(mesh, dcomp, b, ix, iy, f, kind[1])
        return one(comp[i], g, None if dcomp is None else dcomp[i], gd, coef(hc), i)
    # finer: the two fine cells' fluxes across this face
    tot = None
    for close, far, sT in fine_cells(b, ix, iy, f, kind[1:]):
        g = coarse_side_ghost(comp, b, ix, iy, f, close, far, sT)
        gd = None if dcomp is None else coarse_side_ghost(dcomp, b, ix, iy, f, close, far, sT)
        fl = one(comp[close], g, None if dcomp is None else dcomp[close], gd, coef(h[close // 64]),
                 close)
        tot = fl if tot is None else tot + fl
    return tot


def _lab(mesh, f, xs, ys):
    """(n_blocks, 10, 10) array of f with a one-cell halo: neighbours' cells,
    or xs / ys times the boundary cell at domain faces (index [b, y+1, x+1]),
    zero at coarse-fine faces."""
    nb = _neighbours(mesh)
    blk = np.asarray(f, dtype=np.float64).reshape(-1, BS, BS)
    lab = np.zeros((blk.shape[0], BS + 2, BS + 2))
    lab[:, 1:-1, 1:-1] = blk
    for b in range(blk.shape[0]):
        xm, xp, ym, yp = nb[b]
        # a coarse-fine face (-2) keeps a zero ghost: only explicit cells,
        # computed in face form, sit on it
        lab[b, 1:-1, 0] = blk[xm, :, BS - 1] if xm >= 0 else (xs * blk[b, :, 0] if xm == -1 else 0.0)
        lab[b, 1:-1, -1] = blk[xp, :, 0] if xp >= 0 else (xs * blk[b, :, BS - 1] if xp == -1 else 0.0)
        lab[b, 0, 1:-1] = blk[ym, BS - 1, :] if ym >= 0 else (ys * blk[b, 0, :] if ym == -1 else 0.0)
        lab[b, -1, 1:-1] = blk[yp, 0, :] if yp >= 0 else (ys * blk[b, BS - 1, :] if yp == -1 else 0.0)
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
    out = t.reshape(-1)
    cells = explicit_cells(mesh)
    if cells:
        fc = faces(mesh)
        hb = _h(mesh)
        comps = [np.asarray(a, dtype=np.float64).reshape(-1) for a in (u, v, udef_u, udef_v, chi)]
        U, V, UD, VD, CHI = comps
        fd = lambda hh: 0.5 * hh / dt
        for b, ix, iy in cells:
            acc = None
            for f in range(4):
                comp, dcomp = (U, UD) if f < 2 else (V, VD)
                fl = _face_flux(mesh, fc, b, ix, iy, f, hb, fd, comp, dcomp, CHI, -1.0)
                term = -fl if f in (0, 2) else fl
                acc = term if acc is None else acc + term
            out[_gi(b, ix, iy)] = acc
    return out


def pressure_correct(mesh, p, u, v, dt):
    h = _h(mesh)[:, None, None]
    pFac = -0.5 * dt * h
    ih2 = 1.0 / h / h
    dpx, dpy = _diffs(_lab(mesh, p, 1.0, 1.0))
    tu = pFac * dpx
    tv = pFac * dpy
    uo = np.asarray(u, dtype=np.float64).reshape(-1, BS, BS) + tu * ih2
    vo = np.asarray(v, dtype=np.float64).reshape(-1, BS, BS) + tv * ih2
    uo, vo = uo.reshape(-1), vo.reshape(-1)
    cells = explicit_cells(mesh)
    if cells:
        fc = faces(mesh)
        hb = _h(mesh)
        P = np.asarray(p, dtype=np.float64).reshape(-1)
        u0, v0 = np.asarray(u, dtype=np.float64).reshape(-1), np.asarray(v, dtype=np.float64).reshape(-1)
        pf = lambda hh: -0.5 * dt * hh
        for b, ix, iy in cells:
            t = []
            for f in range(4):
                fl = _face_flux(mesh, fc, b, ix, iy, f, hb, pf, P, None, None, 1.0)
                t.append(-fl if f in (0, 2) else fl)
            i = _gi(b, ix, iy)
            ihb = 1.0 / hb[b] / hb[b]
            uo[i] = u0[i] + (t[0] + t[1]) * ihb
            vo[i] = v0[i] + (t[2] + t[3]) * ihb
    return uo, vo


def laplacian_face_form(mesh, p) -> np.ndarray:
    """sum over the faces of (p_nb - p_self) with the coarse-fine ghosts
    above (coarser face: ghost - self; finer face: sum over the two fine
    cells of (fine - fine ghost)), wall faces contributing nothing: the
    operator the explicit matrix assembles (ExpAMRSolver.cpp:131-313), at
    every explicit cell.  Returns {global row: value}."""
    fc = faces(mesh)
    P = np.asarray(p, dtype=np.float64).reshape(-1)
    out = {}
    for b, ix, iy in explicit_cells(mesh):
        i = _gi(b, ix, iy)
        edge = (ix == 0, ix == 7, iy == 0, iy == 7)
        acc = 0.0
        for f, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            if not edge[f]:
                acc += P[_gi(b, ix + dx, iy + dy)] - P[i]
                continue
            kind = fc[b][f]
            if kind[0] == "same":
                jx = {0: 7, 1: 0}.get(f, ix)
                jy = {2: 7, 3: 0}.get(f, iy)
                acc += P[_gi(kind[1], jx, jy)] - P[i]
            elif kind[0] == "coarse":
                acc += fine_ghost(mesh, P, b, ix, iy, f, kind[1]) - P[i]
            elif kind[0] == "fine":
                for close, far, sT in fine_cells(b, ix, iy, f, kind[1:]):
                    acc += P[close] - coarse_side_ghost(P, b, ix, iy, f, close, far, sT)
        out[i] = acc
    return out
