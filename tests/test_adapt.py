"""CPU checks of mesh adaptation: the adapted mesh (cup2d_mesh_adapt) covers
the domain, is 2:1 balanced over faces and corners, honours the tags
(refined where asked below the top level, compressed only whole sibling
groups), and the restated field transfer (oracle/flow.py) is exact for
linear fields, keeps each parent cell's mean one level down and the
integral under compression."""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import workloads as W
from oracle import flow as FL


def _balanced(m):
    where = {(int(l), int(i), int(j)) for l, i, j in zip(m.level, m.bx, m.by)}

    def covering(l, i, j):
        for mm in range(l, -1, -1):
            d = l - mm
            if (mm, i >> d, j >> d) in where:
                return mm
        return -1

    for l, i, j in where:
        nx, ny = m.bpdx << l, m.bpdy << l
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if not di and not dj:
                    continue
                ni, nj = i + di, j + dj
                if not (0 <= ni < nx and 0 <= nj < ny):
                    continue
                c = covering(l, ni, nj)
                if c >= 0 and c < l - 1:
                    return False
    return True


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_adapt_invariants(seed):
    m = W.cylinder_mesh(levels=4, target=600)
    rng = np.random.default_rng(seed)
    tags = rng.choice([-1, 0, 1], size=m.n_blocks, p=[0.6, 0.3, 0.1])
    m2, kind, src = M.adapt(m, tags)
    h = m.h()
    h2 = m2.h()
    assert np.isclose((h * h).sum() * 64, (h2 * h2).sum() * 64, rtol=1e-14)
    assert _balanced(m2)
    old = {(int(l), int(i), int(j)): b for b, (l, i, j) in enumerate(zip(m.level, m.bx, m.by))}
    new = {(int(l), int(i), int(j)) for l, i, j in zip(m2.level, m2.bx, m2.by)}
    for (l, i, j), b in old.items():
        if tags[b] > 0 and l + 1 < m.levels:  # refined: no longer a leaf
            assert (l, i, j) not in new
        if tags[b] >= 0:  # never compressed
            assert (l - 1, i >> 1, j >> 1) not in new or l == 0
    for b, (l, i, j) in enumerate(zip(m2.level, m2.bx, m2.by)):
        if kind[b] == 2:  # whole sibling group, all tagged -1
            kids = [old[(int(l) + 1, 2 * int(i) + (q & 1), 2 * int(j) + (q >> 1))] for q in range(4)]
            assert list(src[b]) == kids and all(tags[k] < 0 for k in kids)
        elif kind[b] == 1:
            a = int(src[b][0])
            d = int(l) - int(m.level[a])
            assert d >= 1 and (int(i) >> d, int(j) >> d) == (int(m.bx[a]), int(m.by[a]))
        else:
            assert (int(l), int(i), int(j)) == (int(m.level[src[b][0]]), int(m.bx[src[b][0]]),
                                                int(m.by[src[b][0]]))
    assert (kind == 1).sum() > 0 and (kind == 2).sum() > 0


def test_transfer_linear_exact_and_conservative():
    m = W.cylinder_mesh(levels=4, target=600)
    rng = np.random.default_rng(5)
    tags = rng.choice([-1, 0, 1], size=m.n_blocks, p=[0.6, 0.2, 0.2])
    m2, kind, src = M.adapt(m, tags)
    x, y = m.cell_centers()
    x2, y2 = m2.cell_centers()
    f = 0.3 - 1.1 * x + 2.4 * y
    g = FL.transfer(m, m2, kind, src, f)
    assert np.allclose(g, 0.3 - 1.1 * x2 + 2.4 * y2, rtol=0, atol=1e-13)
    # a random field: compression keeps the integral of the four children,
    # one-level refinement keeps each parent cell's mean
    r = rng.normal(size=m.n)
    g = FL.transfer(m, m2, kind, src, r)
    w, w2 = m.cell_h() ** 2, m2.cell_h() ** 2
    for b in np.flatnonzero(kind == 2):
        kids = src[b]
        assert np.isclose(np.dot(w2[64 * b:64 * b + 64], g[64 * b:64 * b + 64]),
                          sum(np.dot(w[64 * k:64 * k + 64], r[64 * k:64 * k + 64]) for k in kids),
                          rtol=1e-13, atol=1e-15)
    par = {}
    for b in np.flatnonzero(kind == 1):
        a = int(src[b][0])
        if int(m2.level[b]) - int(m.level[a]) == 1:
            par.setdefault(a, []).append(b)
    full = [a for a, kids in par.items() if len(kids) == 4]
    assert full
    for a in full:
        tot = sum(np.dot(w2[64 * b:64 * b + 64], g[64 * b:64 * b + 64]) for b in par[a])
        assert np.isclose(tot, np.dot(w[64 * a:64 * a + 64], r[64 * a:64 * a + 64]), rtol=1e-12, atol=1e-14)


def test_refine_then_compress_is_identity():
    """Refining every block one level and compressing every block back gives
    the mesh and, to rounding, the field again (the four children of a parent
    cell carry vc +- gx/4 +- gy/4, whose mean is vc)."""
    m = M.uniform(4, 2, 1.0)
    m.levels = 2
    f = np.random.default_rng(7).normal(size=m.n)
    m2, kind, src = M.adapt(m, np.ones(m.n_blocks, np.int8))
    assert m2.n_blocks == 4 * m.n_blocks and (kind == 1).all()
    g = FL.transfer(m, m2, kind, src, f)
    m3, kind3, src3 = M.adapt(m2, -np.ones(m2.n_blocks, np.int8))
    assert (kind3 == 2).all() and m3.n_blocks == m.n_blocks
    assert np.array_equal(m3.level, m.level)
    order = {(int(i), int(j)): b for b, (i, j) in enumerate(zip(m.bx, m.by))}
    perm = np.array([order[(int(i), int(j))] for i, j in zip(m3.bx, m3.by)])
    back = FL.transfer(m2, m3, kind3, src3, g).reshape(-1, 64)
    assert np.allclose(back, f.reshape(-1, 64)[perm], rtol=0, atol=1e-15)


def test_block_score_rotation_and_interface():
    """A solid-body rotation u = -w (y - y0), v = w (x - x0) has vorticity 2 w
    everywhere: every block off the domain walls (whose ghosts are the
    boundary cell's, not the field's) scores 2 |w| across same-level and
    coarse-fine faces; blocks whose cells or 2-3 ghost layers hold the disk's
    interface score 2 rtol, the others keep the vorticity."""
    m = W.cylinder_mesh(levels=4, target=600)
    x, y = m.cell_centers()
    w = 0.7
    u, v = -w * (y - 0.5), w * (x - 0.3)
    z = np.zeros(m.n)
    sc = FL.block_score(m, u, v, z, 5.0)
    nx, ny = m.bpdx << m.level, m.bpdy << m.level
    inner = (m.bx > 0) & (m.by > 0) & (m.bx < nx - 1) & (m.by < ny - 1)
    assert inner.sum() > 100
    assert np.allclose(sc[inner], 2 * w, rtol=1e-12)
    chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
    sc2 = FL.block_score(m, u, v, chi, 5.0)
    C = chi.reshape(-1, 64)
    itf = ((C > 0) & (C < 0.9)).any(axis=1)
    assert (sc2[itf] == 10.0).all()
    hit = sc2 == 10.0
    assert hit.sum() > itf.sum()  # the ghost layers reach further
    assert np.array_equal(sc2[~hit], sc[~hit])


def test_adapt_rejects_bad_input():
    m = M.uniform(4, 2, 1.0)
    with pytest.raises(ValueError):  # one tag per block
        M.adapt(m, np.zeros(m.n_blocks + 1, np.int8))
    # the top level cannot be refined: tags there leave the block as it is
    m2, kind, src = M.adapt(m, np.ones(m.n_blocks, np.int8))
    assert m2.n_blocks == m.n_blocks and (kind == 0).all()
    assert np.array_equal(src[:, 0], np.arange(m.n_blocks)) and (src[:, 1:] == -1).all()
