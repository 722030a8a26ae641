"""Writes cylinder100.npz: 100 time steps of the flow past the cylinder
(r = 0.05 at (0.3, 0.5), U = 0.2, Re = 1000, CFL 0.4, lambda = 1e6) on the
static 5-level mesh (workloads.cylinder_mesh(), 8348 blocks), run on the CPU
with the oracle's restatement of the reference's time step:

    dt from calcMaxTimestep (Simulation.cpp:426-460, no ramp-up)
    advection-diffusion, two stages          (advDiff.cpp:238-282)
    penalization toward the body velocity    (PressureSingle.cpp:268-322)
    rhs = div u, L p_old subtracted, FP64 BiCGSTAB to 1e-12 relative 2-norm,
    p = p_old + dp - <dp>_h2                 (PressureSingle.cpp:827-879)
    u -= dt grad p                           (PressureSingle.cpp:122-210)

The record holds every step's dt and pressure iteration count, and the
per-block means and maxima of |.| of u, v and p after steps 10 and 100
(float64, 8348 per field), so the GPU test can replay the same dt sequence
and compare the fields at full size without storing them whole.

Run from the repository root (about 45 minutes on 8 cores):
    python tests/golden/make_cylinder100.py
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cup2d_amd import workloads as W  # noqa: E402
from oracle import flow as FL  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle import projection as PR  # noqa: E402

STEPS = 100
KEEP = (10, 100)
REL_TOL = 1e-12
LAMBDA = 1e6


def block_stats(f):
    b = f.reshape(-1, 64)
    return b.mean(axis=1), np.abs(b).max(axis=1)


def main():
    m = W.cylinder_mesh()
    uinf = (W.CYL_U, 0.0)
    us, vs = -uinf[0], -uinf[1]
    A = O.assemble(m)
    w = np.repeat(O.h2(m), 64)
    chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
    u = np.zeros(m.n)
    v = np.zeros(m.n)
    p = np.zeros(m.n)
    z = np.zeros(m.n)
    dts, iters = [], []
    out = {"n_blocks": np.int64(m.n_blocks), "rel_tol": np.float64(REL_TOL)}
    t0 = time.time()
    for k in range(1, STEPS + 1):
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
        if k in KEEP:
            for name, f in (("u", u), ("v", v), ("p", p)):
                mean, amax = block_stats(f)
                out[f"{name}_mean_{k}"] = mean
                out[f"{name}_absmax_{k}"] = amax
        print(f"step {k} dt {dt:.6e} iterations {st['iterations']} ({time.time() - t0:.0f} s)", flush=True)
    out["dt"] = np.array(dts)
    out["iterations"] = np.array(iters, dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "cylinder100.npz"), **out)


if __name__ == "__main__":
    main()
