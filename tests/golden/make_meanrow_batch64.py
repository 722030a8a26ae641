"""Writes meanrow_batch64.json: the oracle's outcome on every instance of the
64-instance batch (block counts log-uniform in [512, 8192], seed 4; RHS
vortex_rhs(mesh, 40 + k)) with the mean-constraint row
(bMeanConstraint = 1: the row of cell (0, 0) of block (0, 0, 0) replaced
by sum h^2 x, its RHS 0; ExpAMRSolver.cpp:252-256, :340-345, BiCGSTAB.cu:393-409),
solved by the plain BiCGSTAB restatement (BiCGSTAB.cu:412-620) to 1e-10
relative 2-norm with at most 3000 iterations and 100 restarts.

The mean-row system is much harder for block-Jacobi BiCGSTAB than the
mean-free one: on the CPU in FP64 many instances stall short of 1e-10
within 3000 iterations.  The GPU test compares the batched solver's
outcome with this record.

Run from the repository root (about 15 minutes on 8 cores):
    python tests/golden/make_meanrow_batch64.py
"""

from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cup2d_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

MAX_ITER = 3000


def main():
    out = {"max_iter": MAX_ITER, "rel_tol": 1e-10, "norm": "l2", "max_restarts": 100,
           "rhs_seed_base": 40, "instances": []}
    for k, m in enumerate(W.batch64_meshes()):
        b = W.vortex_rhs(m, 40 + k)
        A = O.assemble(m, 1)
        _, st = O.bicgstab(A, b, mean_mode=1, h2v=O.h2(m), max_iter=MAX_ITER, rel_tol=1e-10,
                           norm="l2", max_restarts=100)
        out["instances"].append({"k": k, "n_blocks": int(m.n_blocks), "iterations": st["iterations"],
                                 "converged": st["converged"],
                                 "best_over_initial": st["err_opt"] / st["err_init"]})
        print(k, m.n_blocks, st["iterations"], st["converged"], flush=True)
    with open(os.path.join(HERE, "meanrow_batch64.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
