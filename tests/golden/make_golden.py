"""Writes the golden fixtures of the Poisson path into this directory.

The reference has no golden vectors for this path (its Python tests never
check a Poisson result), and its solver cannot be built here, so the
fixtures come from the CPU oracle once it is pinned by the closed-form,
known-answer, invariant and dense-solve checks of tests/test_oracle.py.
They freeze the operator and the solutions so that a later change of the
oracle or of the product shows up as a fixture mismatch, and they give the
GPU tests reference solutions without re-running the oracle there.

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cup2d_amd import mesh as M  # noqa: E402
from cup2d_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def tiny_cf_mesh():
    """2x2 level-0 blocks with block (0,1) refined: its four children see the
    coarse block (0,0) across their Ym face (7 blocks)."""
    blocks = [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2), (1, 0, 3), (1, 1, 3)]
    return M.from_blocks(2, 2, 2, 1.0, blocks)


def csr_digest(A):
    h = hashlib.sha256()
    for a in (A.row_ptr.astype(np.int64), A.col.astype(np.int32), A.val.astype(np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def solve_case(m, b, mean_constraint=0):
    A = O.assemble(m, mean_constraint)
    h2 = O.h2(m)
    if mean_constraint:
        b = b.copy()
        b[A.mean_row] = 0.0  # ExpAMRSolver.cpp:340-345
        x, st = O.bicgstab(A, b, rel_tol=1e-14, norm="l2", max_iter=20000, mean_mode=1, h2v=h2)
    else:
        x, st = O.bicgstab(A, b, rel_tol=1e-14, norm="l2", max_iter=20000)
    x = O.mean_remove(h2, x)
    return A, x, st


def mesh_arrays(m, prefix=""):
    return {prefix + "level": m.level.astype(np.int32), prefix + "bx": m.bx.astype(np.int32),
            prefix + "by": m.by.astype(np.int32),
            prefix + "shape": np.array([m.bpdx, m.bpdy, m.levels, int(m.periodic_x),
                                        int(m.periodic_y)], np.int32),
            prefix + "extent": np.array([m.extent], np.float64)}


def main():
    O.build()
    np.save(os.path.join(HERE, "pinv.npy"), O.pinv())

    # tiny coarse-fine mesh: full CSR, rhs and solution
    m = tiny_cf_mesh()
    b = W.vortex_rhs(m, 1)
    A, x, st = solve_case(m, b)
    np.savez_compressed(os.path.join(HERE, "tiny_cf.npz"), **mesh_arrays(m), row_ptr=A.row_ptr,
                        col=A.col, val=A.val, b=b, x=x, iterations=st["iterations"])

    # uniform 16x8 grid (128 blocks): CSR digest, rhs, solutions without and
    # with the mean-constraint row
    m = W.uniform_16x8_mesh()
    b = W.vortex_rhs(m, 1)
    A, x, _ = solve_case(m, b)
    A1, x1, _ = solve_case(m, b, mean_constraint=1)
    np.savez_compressed(os.path.join(HERE, "uniform16x8.npz"), **mesh_arrays(m), b=b, x=x,
                        x_mean_row=x1, nnz=A.nnz, mean_row=A1.mean_row)
    with open(os.path.join(HERE, "uniform16x8.csr.sha256"), "w") as f:
        f.write(csr_digest(A) + "\n")

    # small 3-level graded mesh with coarse-fine rows everywhere along two
    # interfaces: CSR digest, rhs, solution
    m = M.graded(8, 4, 3, 2.0, disks=[(0.5, 0.5, 0.1)], bands=[0.1, 0.05])
    b = W.vortex_rhs(m, 2)
    A, x, _ = solve_case(m, b)
    np.savez_compressed(os.path.join(HERE, "graded3.npz"), **mesh_arrays(m), b=b, x=x, nnz=A.nnz)
    with open(os.path.join(HERE, "graded3.csr.sha256"), "w") as f:
        f.write(csr_digest(A) + "\n")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
