"""GPU tests of the device-side assembly (cup2d_ps_set_mesh_on_device).

The operator the kernels build must be the host assembly's bit for bit (the
host assembly is itself bit-identical to the oracle's CSR, test_abi.py):
same sizes (n, nonzeros, explicit rows, mean row), the same y = A x to the
last bit on random x, and the same solves.  Rejected meshes fail with the
host's message.  The oracle SpMV is the independent check of the values.
"""

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _tiny_cf():
    return M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                        (1, 0, 3), (1, 1, 3)])


def _shuffled(m, seed):
    """Same leaves in another block order: row order, explicit-row order and
    the mean row all change with it."""
    p = np.random.default_rng(seed).permutation(m.n_blocks)
    return M.Mesh(m.bpdx, m.bpdy, m.levels, m.extent, m.level[p].copy(), m.bx[p].copy(),
                  m.by[p].copy(), m.periodic_x, m.periodic_y)


MESHES = {
    "tiny_cf": _tiny_cf,
    "uniform16x8": W.uniform_16x8_mesh,
    "graded": W.graded_disk_mesh,
    "graded_shuffled": lambda: _shuffled(W.graded_disk_mesh(), 1),
    "amr6_small": lambda: W.four_disk_amr6_mesh(target=3000),
    "periodic": lambda: M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)],
                                 [0.3, 0.15, 0.08], True, True),
    "periodic_x_only": lambda: M.graded(8, 8, 4, 1.0, [(0.02, 0.5, 0.05), (0.7, 0.98, 0.03)],
                                        [0.2, 0.1, 0.05], True, False),
    "amr6_1M": W.four_disk_amr6_mesh,
    # one block column: a block is its own periodic neighbour
    "self_periodic": lambda: M.graded(1, 2, 1, 1.0, [], [], True, False),
    "self_periodic_amr": lambda: M.graded(1, 1, 3, 1.0, [(0.3, 0.3, 0.1)], [0.3, 0.15], True, True),
    "periodic_cf_wrap": lambda: M.graded(2, 1, 4, 1.0, [(0.05, 0.5, 0.05)], [0.3, 0.15, 0.08], True, True),
}


def _pair(m, mc=0, **cfg):
    sh = PS.Solver(mean_constraint=mc, **cfg).set_mesh(m)
    sd = PS.Solver(mean_constraint=mc, **cfg).set_mesh(m, on_device=True)
    return sh, sd


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("mc", [0, 1])
def test_device_assembly_equals_host(oracle, name, mc):
    m = MESHES[name]()
    sh, sd = _pair(m, mc)
    assert sd.sizes() == sh.sizes()
    rng = np.random.default_rng(5 + mc)
    for _ in range(2):
        x = rng.normal(size=m.n)
        yd = sd.spmv(x)
        assert np.array_equal(yd, sh.spmv(x)), name
    A = oracle.assemble(m, mc)
    ref = oracle.spmv(A, x, mean_mode=mc, h2v=oracle.h2(m))
    assert np.abs(yd - ref).max() <= 1e-14 * np.abs(ref).max()
    assert sd.sizes()["nnz"] == A.row_ptr[-1]


def test_device_assembly_equals_host_batch_meshes():
    """The 64 batch meshes (512..8192 blocks, one random disk each), half of
    them in a shuffled block order."""
    rng = np.random.default_rng(9)
    for k, m in enumerate(W.batch64_meshes()):
        if k % 2:
            m = _shuffled(m, k)
        sh, sd = _pair(m)
        assert sd.sizes() == sh.sizes(), k
        x = rng.normal(size=m.n)
        assert np.array_equal(sd.spmv(x), sh.spmv(x)), k


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_device_assembled_solve_identical(oracle, precision):
    m = W.graded_disk_mesh()
    sh, sd = _pair(m, precision=precision, norm=PS.NORM_L2, max_iter=4000)
    b = W.vortex_rhs(m, 3)
    xh, sth = sh.solve(b, rel_tol=1e-10)
    xd, std = sd.solve(b, rel_tol=1e-10)
    assert std["converged"] and std["iterations"] == sth["iterations"]
    assert np.array_equal(xd, xh)
    tr = oracle.true_residual(oracle.assemble(m), b, xd)
    assert tr["rel2"] <= 1.05e-10, tr


BAD = {
    # a level-2 block next to a level-0 block
    "unbalanced": lambda: M.from_blocks(2, 1, 3, 2.0, [(0, 0, 0)] + [(2, 4 + a, b) for a in range(4)
                                                                     for b in range(4)]),
    "level": lambda: M.from_blocks(2, 1, 2, 1.0, [(0, 0, 0), (2, 1, 0)]),
    "index": lambda: M.from_blocks(2, 1, 2, 1.0, [(0, 0, 0), (0, 2, 0)]),
    "negative_index": lambda: M.from_blocks(2, 1, 2, 1.0, [(0, 0, 0), (0, 0, -1)]),
    "duplicate": lambda: M.from_blocks(2, 1, 1, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 0, 0)]),
    # the first failing block decides, whatever its kind
    "level_before_duplicate": lambda: M.from_blocks(2, 1, 1, 1.0, [(0, 0, 0), (3, 0, 0), (0, 0, 0)]),
    "duplicate_before_index": lambda: M.from_blocks(2, 1, 1, 1.0, [(0, 1, 0), (0, 1, 0), (0, 9, 0)]),
}


@pytest.mark.parametrize("name", list(BAD))
def test_device_assembly_rejects_like_host(name):
    m = BAD[name]()
    with pytest.raises(_lib.SolverError) as eh:
        PS.Solver().set_mesh(m)
    s = PS.Solver()
    with pytest.raises(_lib.SolverError) as ed:
        s.set_mesh(m, on_device=True)
    assert str(ed.value) == str(eh.value)
    # the handle takes a valid mesh afterwards
    good = _tiny_cf()
    s.set_mesh(good, on_device=True)
    x = np.random.default_rng(1).normal(size=good.n)
    assert np.array_equal(s.spmv(x), PS.Solver().set_mesh(good).spmv(x))


def test_device_assembled_mesh_has_no_csr():
    m = _tiny_cf()
    s = PS.Solver().set_mesh(m, on_device=True)
    x = np.ones(m.n)
    with pytest.raises(_lib.SolverError):
        s.spmv(x, kind=1)
    with pytest.raises(_lib.SolverError, match="no CSR copy"):
        s.csr()
    # switching between the two assemblies on one handle
    s.set_mesh(W.graded_disk_mesh())
    s.set_mesh(m, on_device=True)
    s.set_mesh(m)
    assert np.array_equal(s.spmv(x, kind=1), s.spmv(x, kind=0)) or \
        np.abs(s.spmv(x, kind=1) - s.spmv(x, kind=0)).max() <= 1e-14


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_mesh_changes_on_one_handle(precision):
    """Mesh changes of growing and shrinking size on one handle, alternating
    host and device assembly (the handle reuses its device buffers across
    meshes), give the operators and solves of fresh handles."""
    meshes = [W.graded_disk_mesh(), _tiny_cf(), W.four_disk_amr6_mesh(target=3000),
              W.graded_disk_mesh(), W.uniform_16x8_mesh(), W.four_disk_amr6_mesh(target=3000)]
    rng = np.random.default_rng(4)
    s = PS.Solver(norm=PS.NORM_L2, precision=precision, max_iter=4000)
    for k, m in enumerate(meshes):
        s.set_mesh(m, on_device=bool(k % 2))
        fresh = PS.Solver(norm=PS.NORM_L2, precision=precision, max_iter=4000).set_mesh(m)
        x = rng.normal(size=m.n)
        assert np.array_equal(s.spmv(x), fresh.spmv(x)), k
        b = W.vortex_rhs(m, k)
        x1, st1 = s.solve(b, rel_tol=1e-10)
        x2, st2 = fresh.solve(b, rel_tol=1e-10)
        assert st1["converged"] and np.array_equal(x1, x2), (k, st1, st2)


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
@pytest.mark.parametrize("mc", [0, 1])
def test_batch_device_assembly_equals_host(precision, mc):
    """All instances assembled in one device pipeline: the batch's sizes,
    offsets, y = A x and solves equal the host-assembled batch's bit for bit
    (64 uneven meshes, a third of them in shuffled block order)."""
    meshes = [m if k % 3 else _shuffled(m, k) for k, m in enumerate(W.batch64_meshes())]
    cfg = dict(precision=precision, mean_constraint=mc, norm=PS.NORM_L2, max_iter=3000)
    bh = PS.BatchSolver(**cfg).set_meshes(meshes)
    bd = PS.BatchSolver(**cfg).set_meshes(meshes, on_device=True)
    assert bd.sizes() == bh.sizes()
    assert np.array_equal(bd.offsets, bh.offsets)
    x = np.random.default_rng(3).normal(size=bh.n)
    assert np.array_equal(bd.spmv(x), bh.spmv(x))
    b = np.concatenate([W.vortex_rhs(m, 40 + k) for k, m in enumerate(meshes)])
    xh, sth = bh.solve(b, rel_tol=1e-10)
    xd, std = bd.solve(b, rel_tol=1e-10)
    # the same runs, converged or not (the mean-row form stalls on some
    # instances within the iteration cap, with either assembly)
    assert [(s["iterations"], s["converged"]) for s in std] == [(s["iterations"], s["converged"]) for s in sth]
    assert np.array_equal(xd, xh)


def test_batch_device_assembly_rejects_like_host():
    """The first failing mesh is reported, with the single-mesh message."""
    good = [W.graded_disk_mesh(), _tiny_cf()]
    for bad in ([good[0], BAD["duplicate"](), good[1], BAD["unbalanced"]()],
                [good[1], BAD["unbalanced"](), BAD["level"]()],
                [good[0], good[1], BAD["index"]()]):
        with pytest.raises(_lib.SolverError) as eh:
            PS.BatchSolver().set_meshes(bad)
        with pytest.raises(_lib.SolverError) as ed:
            PS.BatchSolver().set_meshes(bad, on_device=True)
        assert str(ed.value) == str(eh.value)
    # a handle that failed takes a valid batch
    bd = PS.BatchSolver()
    with pytest.raises(_lib.SolverError):
        bd.set_meshes([good[0], BAD["level"]()], on_device=True)
    bd.set_meshes(good, on_device=True)
    bh = PS.BatchSolver().set_meshes(good)
    x = np.random.default_rng(5).normal(size=bh.n)
    assert np.array_equal(bd.spmv(x), bh.spmv(x))


def test_batch_mesh_changes_on_one_handle():
    """Batches of changing size on one handle, alternating host and device
    assembly, with a retuning in between (the handle reuses its device
    buffers): the operators and solves of fresh handles."""
    all_m = W.batch64_meshes()
    sets = [all_m[:8], all_m[8:40], all_m[:3], all_m[40:]]
    rng = np.random.default_rng(11)
    bs = PS.BatchSolver(norm=PS.NORM_L2, max_iter=3000)
    for k, ms in enumerate(sets):
        bs.set_meshes(ms, on_device=bool(k % 2))
        fresh = PS.BatchSolver(norm=PS.NORM_L2, max_iter=3000).set_meshes(ms)
        x = rng.normal(size=fresh.n)
        assert np.array_equal(bs.spmv(x), fresh.spmv(x)), k
        b = np.concatenate([W.vortex_rhs(m, 60 + j) for j, m in enumerate(ms)])
        xb, stb = bs.solve(b, rel_tol=1e-10)
        xf, stf = fresh.solve(b, rel_tol=1e-10)
        assert np.array_equal(xb, xf), k
        if k == 1:
            d_b = bs.device_array(b)
            bs.autotune(d_b)
            _, st = bs.solve(b, rel_tol=1e-10)
            assert all(s["converged"] for s in st)
