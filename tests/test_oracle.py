"""Pins the CPU oracle (oracle/oracle.cpp) against facts that do not come from
the oracle itself: closed forms, a textbook operator built independently with
scipy, the reference notebook's known-answer row, conservation invariants,
dense direct solves and committed golden vectors."""

import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

from cup2d_amd import mesh as M
from cup2d_amd import workloads as W


def dst_inverse_closed_form():
    """-K^-1 for K = 5-point Dirichlet matrix on an 8x8 block, from the DST-I
    eigen-decomposition (independent of any factorisation)."""
    j = np.arange(1, 9)
    S = math.sqrt(2.0 / 9.0) * np.sin(np.outer(j, j) * math.pi / 9.0)
    lam1 = 2.0 - 2.0 * np.cos(j * math.pi / 9.0)
    lam = (lam1[:, None] + lam1[None, :]).reshape(-1)  # index ky*8+kx
    S2 = np.kron(S, S)
    return -(S2 * (1.0 / lam)[None, :]) @ S2


def K_matrix():
    T = 2 * np.eye(8) - np.eye(8, k=1) - np.eye(8, k=-1)
    return np.kron(np.eye(8), T) + np.kron(T, np.eye(8))


def test_pinv_closed_form(oracle):
    P = oracle.pinv()
    ref = dst_inverse_closed_form()
    assert np.abs(P - ref).max() < 5e-15
    assert np.abs(K_matrix() @ P + np.eye(64)).max() < 5e-15
    assert np.abs(P - P.T).max() < 1e-15
    d = np.diag(P)
    assert -0.51 < d.min() < -0.50 and -0.31 < d.max() < -0.30


def test_block_cholesky_apply_is_minus_Kinv(oracle):
    """AMRSolver::getZ applies -K^-1 up to its 1e-10 drop of small factor entries."""
    m = M.graded(4, 2, 2, 2.0, [(0.6, 0.4, 0.05)], [0.3])
    rng = np.random.default_rng(0)
    v = rng.normal(size=m.n)
    z = oracle.block_cholesky_apply(m, v)
    ref = (dst_inverse_closed_form() @ v.reshape(-1, 64).T).T.reshape(-1)
    assert np.abs(z - ref).max() / np.abs(ref).max() < 1e-9


def textbook_neumann(m):
    """5-point FV Laplacian (h^2-scaled) with zero-flux walls on a uniform mesh,
    built cell-by-cell on the global grid, then permuted to block order."""
    nx, ny = m.bpdx * 8, m.bpdy * 8
    gid = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for gy in range(ny):
        for gx in range(nx):
            deg = 0
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                x2, y2 = gx + dx, gy + dy
                if 0 <= x2 < nx and 0 <= y2 < ny:
                    rows.append(gid[gy, gx]); cols.append(gid[y2, x2]); vals.append(1.0)
                    deg += 1
            rows.append(gid[gy, gx]); cols.append(gid[gy, gx]); vals.append(-float(deg))
    L = sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny))
    # block-order index of every global cell
    perm = np.empty(nx * ny, np.int64)
    for b in range(m.n_blocks):
        for iy in range(8):
            for ix in range(8):
                perm[b * 64 + iy * 8 + ix] = gid[m.by[b] * 8 + iy, m.bx[b] * 8 + ix]
    return L[perm][:, perm]


@pytest.mark.parametrize("bp", [(2, 1), (4, 2), (16, 8)])
def test_uniform_is_textbook_laplacian(oracle, bp):
    m = M.uniform(*bp, extent=2.0)
    A = oracle.assemble(m).to_scipy()
    L = textbook_neumann(m)
    assert abs(A - L).max() == 0.0
    assert A.nnz == 5 * m.n - 2 * (bp[0] + bp[1]) * 8


def _tiny_cf_mesh():
    """2x2 level-0 blocks; block (0,1) refined: its children (0,2),(1,2) see
    the coarse block (0,0) across their Ym face."""
    blocks = [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2), (1, 0, 3), (1, 1, 3)]
    return M.from_blocks(2, 2, 2, 1.0, blocks)


def test_known_answer_row_loggeranalyser(oracle):
    """tools/LoggerAnalyser.ipynb:496-497: a fine edge cell with one coarser
    face (central Taylor case) has coefficients {1,-10/3,1,4/5,1/12,1/2,-1/20}.
    The notebook checks them by sum(cols*val) over its signed column list."""
    m = _tiny_cf_mesh()
    A = oracle.assemble(m).to_scipy().tocsr()
    fine = 4  # block (1, 1, 2): index[0] odd -> ix_c = ix/2 + 4
    ix, iy = 4, 0  # ix_c = 6 (central), taylorSign(ix even) = -1
    row = fine * 64 + iy * 8 + ix
    coarse = 0
    roles = {
        "left": fine * 64 + iy * 8 + ix - 1,
        "self": row,
        "right": fine * 64 + iy * 8 + ix + 1,
        "inward": fine * 64 + (iy + 1) * 8 + ix,
        "c-1": coarse * 64 + 7 * 8 + 5,
        "c": coarse * 64 + 7 * 8 + 6,
        "c+1": coarse * 64 + 7 * 8 + 7,
    }
    r = A.getrow(row)
    got = dict(zip(r.indices.tolist(), r.data.tolist()))
    assert len(got) == 7
    vals = [got[roles[k]] for k in ("left", "self", "right", "inward", "c-1", "c", "c+1")]
    nb = [1.0, -3.3333333333333333, 1.0, 0.8, 0.08333333333333333, 0.5, -0.05]
    assert np.allclose(vals, nb, rtol=0, atol=2e-16)
    # the notebook's own check: its signed column ids dotted with the values
    cols_nb = [-517, 518, -519, 526, 4922, -4923, 4924]
    assert abs(np.dot(cols_nb, vals) + 4639.4) < 0.0005
    assert abs(sum(vals)) < 1e-15


def _exact_row(mesh, A, row):
    return {int(c): Fraction(v).limit_denominator(100000) for c, v in
            zip(A.getrow(row).indices, A.getrow(row).data)}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_amr_conservation(oracle, seed):
    """Zero row sums (A 1 = 0) and zero column sums (1^T A = 0: the coarse-fine
    fluxes cancel) on random graded meshes; a wrong sign or index in any
    coarse-fine coefficient breaks one of them."""
    rng = np.random.default_rng(seed)
    disks = [(rng.uniform(0.2, 1.8), rng.uniform(0.2, 0.8), rng.uniform(0.02, 0.1)) for _ in range(3)]
    m = M.graded(8, 4, 4, 2.0, disks, [0.4, 0.2, 0.1])
    assert len(set(m.level.tolist())) >= 3
    A = oracle.assemble(m).to_scipy()
    assert np.abs(np.asarray(A.sum(axis=1))).max() < 1e-13
    assert np.abs(np.asarray(A.sum(axis=0))).max() < 1e-13
    # diagonal strictly negative, every row has the self entry
    assert (A.diagonal() < 0).all()


def test_rational_coefficients(oracle):
    """Every coefficient is a small rational from {1, 2/3, 1/5, 8/15 x D1/D2}."""
    m = M.graded(8, 4, 3, 2.0, [(0.7, 0.5, 0.1)], [0.3, 0.1])
    A = oracle.assemble(m)
    for v in np.unique(A.val):
        f = Fraction(float(v)).limit_denominator(480)
        assert abs(float(f) - v) < 1e-14, v
        assert 480 % f.denominator == 0


def test_consistency_second_order(oracle):
    """A f ~ h^2 Lap f for f satisfying the Neumann walls; uniform rows are
    second-order accurate, coarse-fine rows first-order (in the h^2-scaled
    sense their error is O(h^3) vs O(h^4))."""
    errs = []
    for lv in (3, 4):
        m = M.graded(4, 2, lv, 2.0, [(1.0, 0.5, 0.05)], [0.3, 0.15, 0.08][: lv - 1])
        A = oracle.assemble(m).to_scipy()
        x, y = m.cell_centers()
        h = m.cell_h()
        f = np.cos(math.pi * x / 2.0) * np.cos(math.pi * y)
        lap = -(math.pi ** 2) * (0.25 + 1.0) * f
        e = A @ f - h * h * lap
        errs.append(np.abs(e).max() / np.abs(h * h * lap).max())
    assert errs[0] < 0.2 and errs[1] < errs[0]


def dense_solve(A, b, h2):
    """Bordered dense solve of the singular Neumann system with the
    h^2-weighted mean fixed to zero."""
    n = A.shape[0]
    w = np.repeat(h2, 64)
    Bm = np.zeros((n + 1, n + 1))
    Bm[:n, :n] = A.toarray()
    Bm[:n, n] = w
    Bm[n, :n] = w
    return np.linalg.solve(Bm, np.concatenate([b, [0.0]]))[:n]


@pytest.mark.parametrize("kind", ["uniform", "amr"])
def test_bicgstab_vs_dense(oracle, kind):
    m = M.uniform(2, 1, 2.0) if kind == "uniform" else _tiny_cf_mesh()
    A = oracle.assemble(m)
    b = W.vortex_rhs(m, 5)
    h2 = oracle.h2(m)
    x, st = oracle.bicgstab(A, b, rel_tol=1e-13, norm="l2", max_iter=2000)
    assert st["converged"]
    tr = oracle.true_residual(A, b, x)
    assert tr["rel2"] < 1e-12
    xd = dense_solve(A.to_scipy(), b, h2)
    xm = oracle.mean_remove(h2, x)
    assert np.linalg.norm(xm - xd) / np.linalg.norm(xd) < 1e-9


def test_bicgstab_mean_row_vs_dense(oracle):
    """bMeanConstraint=1: the mean row is replaced by sum h^2 x = 0
    (BiCGSTAB.cu:393-409), so the solution is the mean-free one directly."""
    m = _tiny_cf_mesh()
    A = oracle.assemble(m, 1)
    assert A.mean_row == 0
    h2 = oracle.h2(m)
    b = W.vortex_rhs(m, 6)
    b[A.mean_row] = 0.0
    x, st = oracle.bicgstab(A, b, mean_mode=1, h2v=h2, rel_tol=1e-13, norm="l2", max_iter=3000)
    A0 = oracle.assemble(m, 0).to_scipy()
    Ad = A0.toarray()
    Ad[A.mean_row, :] = np.repeat(h2, 64)
    xd = np.linalg.solve(Ad, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    assert abs(np.dot(np.repeat(h2, 64), x)) < 1e-12 * np.abs(x).max()


def test_pipelined_matches_plain(oracle):
    m = W.uniform_16x8_mesh()
    A = oracle.assemble(m)
    b = W.workload_rhs(m, 1)
    h2 = oracle.h2(m)
    x1, s1 = oracle.bicgstab(A, b, rel_tol=1e-11, norm="l2")
    x2, s2 = oracle.pipelined_bicgstab(A, m, b, rel_tol=1e-11)
    assert s1["converged"] and s2["converged"]
    # both stop on their recurrence residual (1e-11); the true residual drifts
    # from it by rounding, which the OpenMP reductions' thread count varies
    for x in (x1, x2):
        assert oracle.true_residual(A, b, x)["rel2"] < 5e-11
    a, c = oracle.mean_remove(h2, x1), oracle.mean_remove(h2, x2)
    assert np.linalg.norm(a - c) / np.linalg.norm(a) < 1e-7


def test_reference_eps_stagnates(oracle):
    """With the reference's eps = 1e-21 (BiCGSTAB.cu:427) the h^2-scaled system
    stagnates far above 1e-10; with eps = 1e-100 it converges."""
    m = W.uniform_16x8_mesh()
    A = oracle.assemble(m)
    b = W.workload_rhs(m, 1)
    _, s21 = oracle.bicgstab(A, b, rel_tol=1e-10, norm="l2", eps=1e-21)
    _, s100 = oracle.bicgstab(A, b, rel_tol=1e-10, norm="l2", eps=1e-100)
    assert not s21["converged"]
    assert s100["converged"]


def test_mean_remove_closed_form(oracle):
    m = _tiny_cf_mesh()
    h2 = oracle.h2(m)
    x = np.random.default_rng(1).normal(size=m.n) + 3.0
    xm = oracle.mean_remove(h2, x)
    assert abs(np.dot(np.repeat(h2, 64), xm)) < 1e-13
    assert np.allclose(x - xm, (x - xm)[0])


@pytest.mark.parametrize("bp", [(2, 1), (16, 8)])
def test_spmv_equals_textbook_product(oracle, bp):
    """orc_spmv (mode 0) is the product with the independently built textbook
    Laplacian on uniform meshes."""
    m = M.uniform(*bp, extent=2.0)
    A = oracle.assemble(m)
    x = np.random.default_rng(12).normal(size=m.n)
    y = oracle.spmv(A, x)
    ref = textbook_neumann(m) @ x
    assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("which", ["tiny_cf", "graded", "periodic"])
def test_spmv_mean_modes_equal_scipy(oracle, which):
    """orc_spmv in every mean mode against scipy's CSR product of the same
    matrix plus the mean terms built here from the mesh's cell sizes:
    mode 0: y = A x; mode 1 (BiCGSTAB.cu:393-409): the mean row's value is
    sum h^2 x, every other row as mode 0; mode 2 (AMRSolver.h:110-118):
    y_i = (A x)_i + h_i^2 sum h^2 x."""
    if which == "tiny_cf":
        m = _tiny_cf_mesh()
    elif which == "graded":
        m = W.graded_disk_mesh()
    else:
        m = M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)], [0.3, 0.15, 0.08], True, True)
    rng = np.random.default_rng(13)
    x = rng.normal(size=m.n)
    w = m.cell_h() ** 2                      # h^2 of every unknown's block
    h2b = oracle.h2(m)
    assert np.array_equal(np.repeat(h2b, 64), w)
    A0 = oracle.assemble(m, 0)
    ax = A0.to_scipy() @ x
    scale = np.abs(ax).max()
    y0 = oracle.spmv(A0, x)
    assert np.abs(y0 - ax).max() <= 1e-14 * scale
    A1 = oracle.assemble(m, 1)
    assert A1.mean_row >= 0
    y1 = oracle.spmv(A1, x, mean_mode=1, h2v=h2b)
    ref1 = A1.to_scipy() @ x
    ref1[A1.mean_row] = np.dot(w, x)
    assert abs(y1[A1.mean_row] - ref1[A1.mean_row]) <= 1e-14 * np.abs(w * x).sum()
    keep = np.arange(m.n) != A1.mean_row
    assert np.abs(y1[keep] - ax[keep]).max() <= 1e-14 * scale
    y2 = oracle.spmv(A0, x, mean_mode=2, h2v=h2b)
    ref2 = ax + w * np.dot(w, x)
    assert np.abs(y2 - ref2).max() <= 1e-13 * np.abs(ref2).max()
    # the residual routine agrees with b - spmv in each mode
    b = rng.normal(size=m.n)
    for A, mode, y in ((A0, 0, y0), (A1, 1, y1), (A0, 2, y2)):
        tr = oracle.true_residual(A, b, x, mean_mode=mode, h2v=h2b)
        assert abs(tr["r2"] - np.linalg.norm(b - y)) <= 1e-12 * tr["r2"]
        assert abs(tr["rinf"] - np.abs(b - y).max()) <= 1e-12 * tr["rinf"]
