"""Writes cylinder_adapt.npz: 100 time steps of the flow past the cylinder
(as make_cylinder100.py) with the mesh adapted as AdaptTheMesh schedules it
(AdaptTheMesh.cpp:73-77: before each of the first ten steps and then every
20 steps), run on the CPU with the oracle's restatement:

    score per block: max |vorticity|, or 2 Rtol at the obstacle's interface
                     (AdaptTheMesh.cpp:11-71, Helpers.h:103-116)
    tags: refine above Rtol = 5, compress below Ctol = 0.01
                     (launch/launchDiskAMR.sh:51-52)
    new tree: cup2d_mesh_adapt (2:1 balanced), fields moved by the oracle's
                     transfer (refinement from the ancestor cell plus its
                     gradient, compression by the mean of four)
    then the time step of make_cylinder100.py on the new mesh.

The record holds the tags of every adaptation (so the GPU run makes the same
meshes), every step's dt, the final mesh's blocks, and the per-block means
and maxima of |.| of u, v and p after steps 10 and 100.

Run from the repository root (a few minutes on 8 cores):
    python tests/golden/make_cylinder_adapt.py
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cup2d_amd import mesh as M  # noqa: E402
from cup2d_amd import workloads as W  # noqa: E402
from oracle import flow as FL  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import projection as PR  # noqa: E402

STEPS = 100
ADAPT_EVERY = 20
KEEP = (10, 100)
REL_TOL = 1e-12
LAMBDA = 1e6
RTOL, CTOL = 5.0, 0.01


def due(done):
    """AdaptTheMesh's schedule, by the number of steps already taken."""
    return done <= 10 or done % ADAPT_EVERY == 0


def block_stats(f):
    b = f.reshape(-1, 64)
    return b.mean(axis=1), np.abs(b).max(axis=1)


def main():
    m = W.cylinder_mesh()
    uinf = (W.CYL_U, 0.0)
    us, vs = -uinf[0], -uinf[1]
    u = np.zeros(m.n)
    v = np.zeros(m.n)
    p = np.zeros(m.n)
    tags_all, tag_steps, dts, iters, nbs = [], [], [], [], []
    out = {"rel_tol": np.float64(REL_TOL), "adapt_every": np.int64(ADAPT_EVERY),
           "rtol": np.float64(RTOL), "ctol": np.float64(CTOL), "n_blocks_start": np.int64(m.n_blocks)}
    t0 = time.time()
    A = None
    for k in range(1, STEPS + 1):
        if due(k - 1):
            chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
            sc = FL.block_score(m, u, v, chi, RTOL)
            tags = np.where(sc > RTOL, 1, np.where(sc < CTOL, -1, 0)).astype(np.int8)
            tags_all.append(tags)
            tag_steps.append(k - 1)
            m2, kind, src = M.adapt(m, tags)
            u, v, p = (FL.transfer(m, m2, kind, src, f) for f in (u, v, p))
            same = (m2.n_blocks == m.n_blocks and np.array_equal(m2.level, m.level)
                    and np.array_equal(m2.bx, m.bx) and np.array_equal(m2.by, m.by))
            if not same:
                A = None
            m = m2
        if A is None:
            A = O.assemble(m)
            w = np.repeat(O.h2(m), 64)
            chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
            z = np.zeros(m.n)
        dt = FL.time_step(m, u, v, uinf, W.CYL_NU, W.CYL_CFL)
        u, v = FL.advect_diffuse(m, u, v, uinf, W.CYL_NU, dt)
        u, v = FL.penalize(u, v, chi, us, vs, LAMBDA, dt)
        b = PR.pressure_rhs(m, u, v, z, z, chi, dt) - O.spmv(A, p)
        dp, st = O.bicgstab(A, b, rel_tol=REL_TOL, norm="l2", max_iter=8000)
        if not st["converged"]:
            raise RuntimeError(f"step {k}: pressure solve did not converge: {st}")
        p = dp + (p - np.dot(w, dp) / np.sum(w))
        u, v = PR.pressure_correct(m, p, u, v, dt)
        dts.append(dt)
        iters.append(st["iterations"])
        nbs.append(m.n_blocks)
        if k in KEEP:
            out[f"level_{k}"] = m.level.astype(np.int32)
            out[f"bx_{k}"] = m.bx.astype(np.int32)
            out[f"by_{k}"] = m.by.astype(np.int32)
            for name, f in (("u", u), ("v", v), ("p", p)):
                mean, amax = block_stats(f)
                out[f"{name}_mean_{k}"] = mean
                out[f"{name}_absmax_{k}"] = amax
        print(f"step {k} blocks {m.n_blocks} dt {dt:.6e} iterations {st['iterations']} "
              f"({time.time() - t0:.0f} s)", flush=True)
    out["tags"] = np.concatenate(tags_all)
    out["tag_offsets"] = np.cumsum([0] + [len(t) for t in tags_all]).astype(np.int64)
    out["tag_steps"] = np.array(tag_steps, dtype=np.int64)
    out["dt"] = np.array(dts)
    out["iterations"] = np.array(iters, dtype=np.int64)
    out["n_blocks"] = np.array(nbs, dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "cylinder_adapt.npz"), **out)


if __name__ == "__main__":
    main()
