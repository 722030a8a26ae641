"""Block-quadtree generator: coverage, 2:1 balance, ordering, config sizes."""

import numpy as np
import pytest

from cup2d_amd import mesh as M
from cup2d_amd import workloads as W


def _cells_at_finest(m):
    L = m.levels - 1
    nx, ny = m.bpdx << L, m.bpdy << L
    owner = -np.ones((ny, nx), np.int64)
    for b in range(m.n_blocks):
        d = L - m.level[b]
        x0, y0 = m.bx[b] << d, m.by[b] << d
        assert (owner[y0:y0 + (1 << d), x0:x0 + (1 << d)] == -1).all(), "overlap"
        owner[y0:y0 + (1 << d), x0:x0 + (1 << d)] = b
    return owner


@pytest.mark.parametrize("seed", [0, 1])
def test_cover_and_balance(seed):
    rng = np.random.default_rng(seed)
    disks = [(rng.uniform(0.2, 1.8), rng.uniform(0.2, 0.8), rng.uniform(0.02, 0.08)) for _ in range(3)]
    m = M.graded(8, 4, 5, 2.0, disks, [0.3, 0.15, 0.08, 0.04])
    owner = _cells_at_finest(m)
    assert (owner >= 0).all()
    lev = m.level[owner]
    # face and corner neighbours differ by at most one level
    for dy, dx in ((0, 1), (1, 0), (1, 1), (1, -1)):
        a = lev[max(dy, 0):lev.shape[0] + min(dy, 0), max(dx, 0):lev.shape[1] + min(dx, 0)]
        b = lev[max(-dy, 0):lev.shape[0] + min(-dy, 0), max(-dx, 0):lev.shape[1] + min(-dx, 0)]
        assert np.abs(a.astype(int) - b).max() <= 1


def test_uniform_counts():
    m = W.uniform_16x8_mesh()
    assert m.n_blocks == 128 and m.n == 8192
    assert np.allclose(m.h(), 1.0 / 64)


def test_config_sizes():
    m2 = W.graded_disk_mesh()
    assert 1300 <= m2.n_blocks <= 1700
    assert set(m2.level.tolist()) == {0, 1, 2} or set(m2.level.tolist()) == {1, 2}


def test_hilbert_locality():
    """Consecutive blocks are spatial neighbours far more often than not."""
    m = W.graded_disk_mesh()
    h = m.h() * 8
    cx = (m.bx + 0.5) * h
    cy = (m.by + 0.5) * h
    d = np.hypot(np.diff(cx), np.diff(cy))
    assert np.median(d / np.maximum(h[1:], h[:-1])) <= 1.01
