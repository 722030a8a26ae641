"""The batched workload at full size: 64 independent AMR instances, block
counts log-uniform in [512, 8192] (216,896 blocks, 13,881,344 unknowns in
all), each with its own disk refinement and vortex right-hand side.

Bar: every instance converges, and its true residual -- recomputed by the
oracle in long double from the returned solution -- is within 1.05e-10
relative (2-norm) in both precisions.
"""

import numpy as np
import pytest

from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    meshes = W.batch64_meshes()
    rhs = [W.workload_rhs(m, 400 + k) for k, m in enumerate(meshes)]
    return meshes, rhs


@pytest.fixture(scope="module")
def c4_ops(oracle, c4):
    return [oracle.assemble(m) for m in c4[0]]


@pytest.mark.parametrize("precision", [PS.FP64, PS.MIXED])
def test_full_batch_converges_to_fp64_residual(oracle, c4, c4_ops, precision):
    meshes, rhs = c4
    bs = PS.BatchSolver(PS.default_config(precision=precision, norm=PS.NORM_L2, max_iter=4000))
    bs.set_meshes(meshes, on_device=True)
    x, st = bs.solve(np.concatenate(rhs), rel_tol=1e-10)
    worst = 0.0
    for k, (A, b) in enumerate(zip(c4_ops, rhs)):
        xk = bs.split(x)[k]
        tr = oracle.true_residual(A, b, xk)
        worst = max(worst, tr["rel2"])
        assert st[k]["converged"], (k, meshes[k].n_blocks, st[k])
        assert tr["rel2"] <= 1.05e-10, (k, meshes[k].n_blocks, tr)
        # the returned solution is mean-free (ExpAMRSolver.cpp:383-415)
        h2 = np.repeat(oracle.h2(meshes[k]), 64)
        assert abs(np.dot(h2, xk)) <= 1e-12 * np.dot(h2, np.abs(xk))
    print(f"precision {precision}: worst true rel2 {worst:.3e}, "
          f"iterations max {max(s['iterations'] for s in st)}")
