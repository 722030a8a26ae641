"""GPU tests of mesh adaptation (row F3): the per-block refinement score and
the field transfer against the oracle's restatement (oracle/flow.py), bit
for bit, and the cylinder flow adapting its mesh every step against the same
steps on the CPU (the same tags, the oracle's transfer and operators)."""

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import flow as FLOW
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W
from oracle import flow as FL
from oracle import projection as PR

pytestmark = pytest.mark.gpu


def _fields(m, seed):
    rng = np.random.default_rng(seed)
    x, y = m.cell_centers()
    u = 0.3 * np.sin(9 * x) * np.cos(7 * y) + 0.01 * rng.normal(size=m.n)
    v = -0.2 * np.cos(8 * x) * np.sin(9 * y) + 0.01 * rng.normal(size=m.n)
    return u, v


@pytest.mark.parametrize("on_device", [False, True])
def test_block_score_matches_oracle(on_device):
    m = W.cylinder_mesh(levels=4, target=600)
    s = PS.Solver().set_mesh(m, on_device=on_device)
    u, v = _fields(m, 1)
    chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
    d = [PS.DeviceArray.from_host(s, a) for a in (u, v, chi)]
    sc = s.block_score_device(*d, 5.0)
    ref = FL.block_score(m, u, v, chi, 5.0)
    assert np.array_equal(sc, ref)
    assert (sc == 10.0).sum() > 10 and (sc < 10.0).sum() > 100


def test_transfer_matches_oracle():
    m = W.cylinder_mesh(levels=4, target=600)
    rng = np.random.default_rng(2)
    tags = rng.choice([-1, 0, 1], size=m.n_blocks, p=[0.6, 0.2, 0.2])
    m2, kind, src = M.adapt(m, tags)
    s = PS.Solver().set_mesh(m)
    f = [rng.normal(size=m.n) for _ in range(3)]
    d_old = [PS.DeviceArray.from_host(s, a) for a in f]
    d_new = [PS.DeviceArray(s, m2.n) for _ in f]
    s.transfer_device(m, m2, kind, src, d_old, d_new)
    for a, dn in zip(f, d_new):
        assert np.array_equal(dn.numpy(), FL.transfer(m, m2, kind, src, a))
    assert (kind == 1).any() and (kind == 2).any()
    # plans that do not match the meshes, aliased fields
    with pytest.raises(ValueError):  # sizes checked by the Python layer
        s.transfer_device(m, m2, kind, src, d_old, d_old)
    m3, k3, s3 = M.adapt(m, np.zeros(m.n_blocks, np.int8))  # the same mesh
    assert m3.n == m.n and (k3 == 0).all()
    with pytest.raises(_lib.SolverError):  # the library's alias check
        s.transfer_device(m, m3, k3, s3, d_old, d_old)
    bad = src.copy()
    b2 = int(np.flatnonzero(kind == 2)[0])
    bad[b2, [0, 1]] = bad[b2, [1, 0]]
    with pytest.raises(_lib.SolverError):
        s.transfer_device(m, m2, kind, bad, d_old, d_new)
    bad = src.copy()
    b0 = int(np.flatnonzero(kind == 0)[0])
    bad[b0, 0] = (bad[b0, 0] + 1) % m.n_blocks
    with pytest.raises(_lib.SolverError):
        s.transfer_device(m, m2, kind, bad, d_old, d_new)


def test_adapting_cylinder_matches_cpu_run(oracle):
    """Four steps of the cylinder flow adapting its mesh before every step
    (AdaptTheMesh's schedule for the first ten steps), FP64: the same tags on
    both sides; velocity and pressure agree with the CPU run to 1e-7 and the
    oracle's own scores give the same tags."""
    m = W.cylinder_mesh(levels=3, target=300)
    flow = FLOW.CylinderFlow(m, precision=PS.FP64, rel_tol=1e-12, max_iter=6000, adapt_every=20)
    u = np.zeros(m.n)
    v = np.zeros(m.n)
    p = np.zeros(m.n)
    cm = m
    changed = 0
    for k in range(4):
        chi = FL.chi_disk(cm, *W.CYL_CENTER, W.CYL_RADIUS)
        ref_tags = np.where((sc := FL.block_score(cm, u, v, chi, flow.rtol)) > flow.rtol, 1,
                            np.where(sc < flow.ctol, -1, 0))
        # the GPU adapts first inside step(); the CPU follows with its tags
        flow.adapt_every = 20
        tags = flow.adapt()
        assert np.mean(tags == ref_tags) >= 0.99, k
        m2, kind, src = M.adapt(cm, tags)
        changed += int(m2.n_blocks != cm.n_blocks)
        u, v, p = (FL.transfer(cm, m2, kind, src, a) for a in (u, v, p))
        cm = m2
        assert flow.mesh.n_blocks == cm.n_blocks
        chi = FL.chi_disk(cm, *W.CYL_CENTER, W.CYL_RADIUS)
        dt = FL.time_step(cm, u, v, flow.uinf, flow.nu, flow.cfl)
        flow.adapt_every = 0  # adapted above
        st = flow.step(dt)
        assert st["converged"], (k, st)
        A = oracle.assemble(cm)
        w = np.repeat(oracle.h2(cm), 64)
        u, v = FL.advect_diffuse(cm, u, v, flow.uinf, flow.nu, dt)
        u, v = FL.penalize(u, v, chi, flow.us, flow.vs, flow.lam, dt)
        z = np.zeros(cm.n)
        b = PR.pressure_rhs(cm, u, v, z, z, chi, dt) - oracle.spmv(A, p)
        dp, _ = oracle.bicgstab(A, b, rel_tol=1e-12, norm="l2", max_iter=6000)
        p = dp + (p - np.dot(w, dp) / np.sum(w))
        u, v = PR.pressure_correct(cm, p, u, v, dt)
    assert changed >= 1
    gu, gv, gp = flow.fields()
    for g, r in ((gu, u), (gv, v), (gp, p)):
        assert np.linalg.norm(g - r) <= 1e-7 * np.linalg.norm(r)
