"""Benchmark of the pressure-Poisson hot path on one MI355X.

One step = one solve A x = b to 1e-10 relative residual (2-norm, x0 = 0) on
the 3-level graded AMR mesh (16x8 base blocks of 8x8 cells, disk at
(0.5,0.5) r=0.1, 1448 blocks / 92,672 unknowns), FP64, with b and x resident
in HBM when the timed region starts.  Prints one JSON line.

Also measured in the same run:
* the same solve in mixed precision;
* the batched workload: 64 independent AMR instances (block counts
  log-uniform in [512, 8192], 216,896 blocks / 13.9M unknowns in total),
  solved together in FP64 and in mixed precision, against the same 64
  solves one after another with the single-solve FP64 solver;
* roofline of the operator kernel (timed live with HIP events on the
  solver's stream) on the bench mesh and on the DRAM-resident batch,
  against the 8 TB/s HBM peak;
* the CPU baseline: the oracle's restatement of the reference CPU solver
  (pipelined BiCGSTAB with the block-Cholesky preconditioner,
  AMRSolver.cpp:149-490) on a bounded sample, OpenMP on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Poisson solves/sec at 1e-10 rel residual, FP64 vs mixed; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


KERNELS = ["p update + block preconditioner", "nu = A z1 (+ dots)", "s update + block preconditioner",
           "t = A z2 (+ dots)", "x, r update (+ dots)"]


def kernel_bytes(which, n, n_blocks, fp64=True):
    """Algorithmic HBM bytes of one launch of iteration kernel `which`: the
    vectors it must read and write once (s = 8 B FP64 / 4 B FP32 Krylov
    vectors, x and x_opt always 8 B), plus 32 B of operator metadata per
    block for the operator kernels (explicit rows add < 1 %)."""
    s = 8 if fp64 else 4
    per_cell = {0: 5 * s + (0 if fp64 else 4),   # r, p, nu in; p, z1 out (+ projected r)
                1: 3 * s, 2: 4 * s, 3: 3 * s,    # z, aux in; y out / r, nu in; r, z2 out
                4: 24 + 6 * s}[which]            # x in/out, x_opt; z1, z2, r, t, rhat in; r out
    meta = 32 * n_blocks if which in (1, 3) else 0
    return per_cell * n + meta


def pmc_traffic(mode):
    """HBM bytes per launch of the kernel of `--kernel-only MODE`, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/
    pmc_traffic.txt), each scaled by its plain-copy calibration factor;
    None when the summary has no line for it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    rows, ff, wf = {}, None, None
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        if f[0] == "fetch_factor":
            ff = float(f[1])
        elif f[0] == "write_factor":
            wf = float(f[1])
        elif len(f) >= 4:
            rows[f[0]] = (float(f[2]), float(f[3]))
    if mode not in rows or ff is None or wf is None:
        return None
    fetch, write = rows[mode]
    return (fetch * ff + write * wf) * 1024.0


def spmv_bytes_csr(n, nnz):
    """Algorithmic bytes of one FP64 CSR SpMV with int32 indices:
    nnz*(8+4) + (n+1)*4 + n*8 (x) + n*8 (y)."""
    return nnz * 12 + (n + 1) * 4 + n * 16


def cpu_baseline(mesh, b, budget_s=12.0):
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    O.set_num_threads(threads)
    A = O.assemble(mesh)
    h2 = O.h2(mesh)
    solves = 0
    t0 = time.perf_counter()
    iters = []
    while True:
        x, st = O.pipelined_bicgstab(A, mesh, b, rel_tol=1e-10, max_iter=2000, max_restarts=100,
                                     h2v=h2)
        solves += 1
        iters.append(st["iterations"])
        if time.perf_counter() - t0 > budget_s or solves >= 20:
            break
    dt = time.perf_counter() - t0
    return {"value": solves / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{solves} full solves of the bench workload (pipelined BiCGSTAB + block "
                      f"Cholesky, AMRSolver.cpp:149-490, to 1e-10 rel 2-norm, {iters[0]} iterations "
                      f"each) in {dt:.1f} s on {threads} OpenMP threads"}


def batch_spmv_handle(PS, meshes):
    """FP64 batch handle with 32 blocks per work item (the tuned size of the
    FP64 batch) for the operator kernel alone."""
    cfg = PS.default_config(norm=PS.NORM_L2)
    cfg.reserved_i[1] = 32
    return PS.BatchSolver(cfg).set_meshes(meshes)


def batch_spmv_roofline(PS, meshes, reps):
    bs = batch_spmv_handle(PS, meshes)
    sz = bs.sizes()
    d_x = bs.device_array(np.random.default_rng(0).normal(size=bs.n))
    d_y = bs.device_array()
    ms = bs.spmv_time(d_x, d_y, reps=reps)
    ms_copy = bs.copy_time(d_x, d_y, reps=reps)
    byt = spmv_bytes_csr(sz["n"], sz["nnz"])
    ach = byt / (ms * 1e-3) / 1e9
    traffic = pmc_traffic("batch_spmv")
    return {"bound": "hbm", "kernel": "kb_spmv (batched stencil SpMV, FP64, 13.9M rows)",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic, "algorithmic_bytes_csr_equivalent": byt, "ms_per_launch": ms,
            # measured DRAM bytes (calibrated counters) over the kernel time, and the
            # streaming-copy ceiling of the same access pattern on this GPU
            "dram_GB_per_s": traffic / (ms * 1e-3) / 1e9 if traffic else None,
            "stream_copy_GB_per_s": 16 * sz["n"] / (ms_copy * 1e-3) / 1e9}


def bench_amr6(PS, W, steps, use_graph):
    """Mixed precision at scale: FP64 and mixed single solves on the 6-level
    AMR mesh around 4 seeded disks (~16k blocks, ~1M unknowns), to 1e-10
    relative 2-norm residual, and the relative L2 difference of the mixed
    pressure from the FP64 one (same mean removal)."""
    mesh = W.four_disk_amr6_mesh()
    b = W.workload_rhs(mesh, 3)
    out = {"workload": "6-level AMR grid refined around 4 seeded disks (r in [0.02, 0.08]), "
                       "seeded vortex RHS, to 1e-10 rel 2-norm",
           "n_blocks": mesh.n_blocks, "n": mesh.n}
    xs = {}
    for name, prec in (("fp64", PS.FP64), ("mixed", PS.MIXED)):
        s = PS.Solver(device=0, norm=PS.NORM_L2, max_iter=6000, use_graph=use_graph,
                      precision=prec)
        t0 = time.perf_counter()
        s.set_mesh(mesh)
        setup_ms = (time.perf_counter() - t0) * 1e3
        d_b = PS.DeviceArray.from_host(s, b)
        d_x = PS.DeviceArray(s, mesh.n)
        zeros = np.zeros(mesh.n)
        d_x.upload(zeros)
        s.solve_device(d_b, d_x, rel_tol=1e-10)  # warm-up
        t = []
        for _ in range(steps):
            d_x.upload(zeros)
            t0 = time.perf_counter()
            st = s.solve_device(d_b, d_x, rel_tol=1e-10)
            t.append(time.perf_counter() - t0)
        sec = float(np.median(t))
        xs[name] = d_x.numpy()
        out[name] = {"solves_per_s": 1.0 / sec, "ms_per_solve": sec * 1e3, "set_mesh_ms": setup_ms,
                     "iterations": st["iterations"], "refinements": st["refinements"],
                     "restarts": st["restarts"], "converged": st["converged"],
                     "true_rel2_residual": s.residual_device(d_b, d_x)["rel2"]}
        del s, d_b, d_x
    out["mixed_speedup_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["mixed_rel_l2_vs_fp64"] = float(np.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                        np.linalg.norm(xs["fp64"]))
    out["assembly"] = bench_assembly(PS, mesh)
    return out


def bench_batch_assembly(PS, meshes, reps=3):
    """Mesh changes of the 64-instance batch: host assembly on threads +
    upload against the device-side assembly of all instances (median of
    `reps` after one warm-up each)."""
    bs = PS.BatchSolver(PS.default_config(norm=PS.NORM_L2))
    res = {}
    for key, dev in (("host_set_meshes_ms", False), ("device_set_meshes_ms", True)):
        bs.set_meshes(meshes, on_device=dev)
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            bs.set_meshes(meshes, on_device=dev)
            t.append(time.perf_counter() - t0)
        res[key] = float(np.median(t)) * 1e3
    res["speedup"] = res["host_set_meshes_ms"] / res["device_set_meshes_ms"]
    return res


def bench_assembly(PS, mesh, reps=5):
    """A mesh change on the host (assembly + upload) against the device-side
    assembly, each the median of `reps` after one warm-up, and whether the two
    operators give bit-identical y = A x."""
    s = PS.Solver(device=0)
    res = {}
    for key, dev in (("host_set_mesh_ms", False), ("device_set_mesh_ms", True)):
        s.set_mesh(mesh, on_device=dev)
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            s.set_mesh(mesh, on_device=dev)
            t.append(time.perf_counter() - t0)
        res[key] = float(np.median(t)) * 1e3
    x = np.random.default_rng(0).normal(size=mesh.n)
    y_dev = s.spmv(x)
    s.set_mesh(mesh)
    res["identical_spmv"] = bool(np.array_equal(y_dev, s.spmv(x)))
    res["speedup"] = res["host_set_mesh_ms"] / res["device_set_mesh_ms"]
    return res


def bench_batch(PS, W, steps, use_graph):
    """The 64-instance batch: batched FP64 and mixed solves/s, the same 64
    solves one after another with the single-solve FP64 solver, and the
    roofline of the batched operator kernel (13.9M unknowns, DRAM-resident)."""
    meshes = W.batch64_meshes()
    rhs = [W.workload_rhs(m, 400 + k) for k, m in enumerate(meshes)]
    b = np.concatenate(rhs)
    out = {"workload": "64 AMR instances, N_b log-uniform in [512, 8192] (seed 4), 5 levels, one "
                       "random disk each, seeded vortex RHS; all solved to 1e-10 rel 2-norm",
           "instances": len(meshes), "n_blocks": int(sum(m.n_blocks for m in meshes))}
    roof = None
    for name, prec in (("fp64", PS.FP64), ("mixed", PS.MIXED)):
        bs = PS.BatchSolver(PS.default_config(precision=prec, norm=PS.NORM_L2, max_iter=4000,
                                              use_graph=use_graph))
        t0 = time.perf_counter()
        bs.set_meshes(meshes)  # host assembly (threads) + upload
        setup_ms = (time.perf_counter() - t0) * 1e3
        d_b = bs.device_array(b)
        d_x = bs.device_array(np.zeros_like(b))
        ib, ms_it = bs.autotune(d_b)
        zeros = np.zeros_like(b)
        d_x.upload(zeros)
        bs.solve_device(d_b, d_x, rel_tol=1e-10)  # warm-up
        t = []
        for _ in range(steps):
            d_x.upload(zeros)
            t0 = time.perf_counter()
            st = bs.solve_device(d_b, d_x, rel_tol=1e-10)
            t.append(time.perf_counter() - t0)
        its = [s["iterations"] for s in st]
        sec = float(np.median(t))
        out[name] = {"solves_per_s": len(meshes) / sec, "ms_per_batch": sec * 1e3,
                     "set_meshes_ms": setup_ms,
                     "item_blocks": ib, "ms_per_iteration_tuned": ms_it,
                     "iterations_max": int(max(its)), "iterations_mean": float(np.mean(its)),
                     "refinements_max": int(max(s["refinements"] for s in st)),
                     "all_converged": bool(all(s["converged"] for s in st))}
        if prec == PS.FP64:
            roof = batch_spmv_roofline(PS, meshes, reps=20)
        del bs, d_b, d_x
    out["assembly"] = bench_batch_assembly(PS, meshes)
    # the same 64 solves one after another (single-solve FP64 solver)
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000, use_graph=use_graph)
    tot = 0.0
    for m, bb in zip(meshes, rhs):
        s.set_mesh(m)
        d_b = PS.DeviceArray.from_host(s, bb)
        d_x = PS.DeviceArray(s, m.n)
        d_x.upload(np.zeros(m.n))
        s.solve_device(d_b, d_x, rel_tol=1e-10)  # warm-up (graph capture)
        d_x.upload(np.zeros(m.n))
        t0 = time.perf_counter()
        s.solve_device(d_b, d_x, rel_tol=1e-10)
        tot += time.perf_counter() - t0
    out["fp64_sequential"] = {"solves_per_s": len(meshes) / tot, "ms_total": tot * 1e3}
    out["mixed_batched_vs_fp64_sequential"] = out["mixed"]["solves_per_s"] / out["fp64_sequential"]["solves_per_s"]
    out["fp64_batched_vs_fp64_sequential"] = out["fp64"]["solves_per_s"] / out["fp64_sequential"]["solves_per_s"]
    return out, roof


def kernel_only(PS, W, which, reps):
    """Launches one kernel `reps` times (for rocprofv3 counter runs) and
    prints its mean time and algorithmic bytes."""
    if which == "projection":
        # the operators either side of the pressure solve on a uniform
        # 128x128-block mesh (1,048,576 cells, walls): algorithmic bytes per
        # cell 48 (rhs: u, v, udef_u, udef_v, chi in, rhs out) and 40
        # (correction: p, u, v in, u, v out)
        from cup2d_amd import mesh as M
        m = M.uniform(128, 128, 1.0)
        s = PS.Solver().set_mesh(m, on_device=True)
        rng = np.random.default_rng(0)
        d = [PS.DeviceArray.from_host(s, rng.normal(size=m.n)) for _ in range(8)]
        out = {}
        for name, fn, byt in (("pressure_rhs", lambda: s.pressure_rhs_device(d[0], d[1], d[2], d[3], d[4], 0.01, d[5]), 48),
                              ("pressure_correct", lambda: s.pressure_correct_device(d[6], 0.01, d[0], d[1]), 40)):
            fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            out[name] = {"ms_per_call": ms, "GB_per_s": byt * m.n / (ms * 1e-3) / 1e9}
        print(json.dumps({"kernel": "projection", "cells": m.n, **out}))
        return
    if which == "assembly":
        # mesh changes on the ~16k-block 6-level mesh: host assembly + upload
        # against the device-side assembly
        print(json.dumps({"kernel": "assembly", **bench_assembly(PS, W.four_disk_amr6_mesh(), reps),
                          "batch": bench_batch_assembly(PS, W.batch64_meshes())}))
        return
    if which == "copy":
        # the library's streaming copy over the batch's 13,881,344 unknowns
        # (8-byte loads, the access pattern of the solver's vector kernels):
        # the bandwidth ceiling and the traffic-counter calibration
        s = batch_spmv_handle(PS, W.batch64_meshes())
        d_x = s.device_array(np.ones(s.n))
        d_y = s.device_array()
        ms = s.copy_time(d_x, d_y, reps=reps)
        print(json.dumps({"kernel": "kb_copy", "ms_per_launch": ms, "launches": reps + 1,
                          "bytes_read": s.n * 8, "bytes_written": s.n * 8,
                          "GB_per_s": 16 * s.n / (ms * 1e-3) / 1e9}))
        return
    if which in ("batch_fp64", "batch_mixed"):
        # 64 iterations of the batched solve (every instance running), plain
        # stream launches so that each kernel is traced: the batched iteration
        # kernels at their real shape
        meshes = W.batch64_meshes()
        b = np.concatenate([W.workload_rhs(m, 400 + k) for k, m in enumerate(meshes)])
        cfg = PS.default_config(precision=PS.FP64 if which == "batch_fp64" else PS.MIXED,
                                max_iter=64, norm=PS.NORM_L2, use_graph=0)
        cfg.reserved_i[1] = 32
        bs = PS.BatchSolver(cfg).set_meshes(meshes)
        d_b = bs.device_array(b)
        d_x = bs.device_array(np.zeros_like(b))
        st = bs.solve_device(d_b, d_x, rel_tol=1e-14)
        print(json.dumps({"kernel": which, "iterations": max(s["iterations"] for s in st),
                          "ms": st[0]["ms"]}))
        return
    if which.startswith("iter"):
        mesh = W.graded_disk_mesh()
        s = PS.Solver(norm=PS.NORM_L2)
        s.set_mesh(mesh)
        k = int(which[4:])
        ms = s.kernel_time(k, reps=reps)
        print(json.dumps({"kernel": KERNELS[k], "ms_per_launch": ms, "launches": reps + 2,
                          "algorithmic_bytes": kernel_bytes(k, mesh.n, mesh.n_blocks)}))
        return
    if which == "spmv":
        mesh = W.graded_disk_mesh()
        s = PS.Solver(norm=PS.NORM_L2)
        s.set_mesh(mesh)
        sz = s.sizes()
        d_x = PS.DeviceArray.from_host(s, np.random.default_rng(0).normal(size=mesh.n))
        d_y = PS.DeviceArray(s, mesh.n)
        ms = s.spmv_time(d_x, d_y, kind=0, reps=reps)
    else:
        meshes = W.batch64_meshes()
        s = batch_spmv_handle(PS, meshes)
        sz = s.sizes()
        d_x = s.device_array(np.random.default_rng(0).normal(size=s.n))
        d_y = s.device_array()
        ms = s.spmv_time(d_x, d_y, reps=reps)
    print(json.dumps({"kernel": which, "ms_per_launch": ms, "launches": reps, "n": sz["n"],
                      "nnz": sz["nnz"], "algorithmic_bytes": spmv_bytes_csr(sz["n"], sz["nnz"])}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--spmv-reps", type=int, default=200)
    ap.add_argument("--graph", type=int, default=1,
                    help="1: replay iteration chunks from a hipGraph; 0: plain stream launches")
    ap.add_argument("--chunk", type=int, default=0, help="iterations per chunk (0: automatic)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--kernel-only", choices=["spmv", "batch_spmv", "batch_fp64", "batch_mixed", "copy", "assembly", "projection", "iter0", "iter1",
                                              "iter2", "iter3", "iter4"], default=None,
                    help="profiling aid: launch only this kernel --steps times and exit "
                         "(spmv: operator on the bench mesh; batch_spmv: batched operator on the "
                         "64-instance batch; batch_fp64 / batch_mixed: 64 iterations of the "
                         "batched solve, every kernel launched from the stream; iterK: iteration "
                         "kernel K of the bench solve; copy: the library's streaming copy over "
                         "13,881,344 doubles, the bandwidth ceiling and the traffic counters' "
                         "calibration; assembly: host against device-side assembly of the "
                         "6-level 16k-block mesh)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the solver has no "
                                   "multi-GPU path in this build", "n_gpus": args.gpus}))
        sys.exit(2)

    from cup2d_amd import poisson as PS
    from cup2d_amd import workloads as W

    if args.kernel_only:
        kernel_only(PS, W, args.kernel_only, args.steps)
        return

    mesh = W.graded_disk_mesh()
    b = W.workload_rhs(mesh, 2, with_sources=True)

    def run_variant(precision, steps, warmup):
        s = PS.Solver(device=0, norm=PS.NORM_L2, max_iter=4000, use_graph=args.graph,
                      chunk=args.chunk, precision=precision)
        t0 = time.perf_counter()
        s.set_mesh(mesh)  # host assembly + upload, reported apart from the solves
        setup_ms = (time.perf_counter() - t0) * 1e3
        d_b = PS.DeviceArray.from_host(s, b)
        d_x = PS.DeviceArray(s, mesh.n)
        zeros = np.zeros(mesh.n)
        for _ in range(warmup):
            d_x.upload(zeros)
            s.solve_device(d_b, d_x, rel_tol=1e-10, max_restarts=100)
        # timed region: `steps` solves, each from x0 = 0
        stats = []
        t0 = time.perf_counter()
        for _ in range(steps):
            d_x.upload(zeros)
            stats.append(s.solve_device(d_b, d_x, rel_tol=1e-10, max_restarts=100))
        t1 = time.perf_counter()
        res = s.residual_device(d_b, d_x)
        return s, d_x, {
            "ms_per_step": (t1 - t0) * 1e3 / steps,
            "solves_per_s": steps / (t1 - t0),
            "iterations": stats[-1]["iterations"],
            "refinements": stats[-1]["refinements"],
            "restarts": stats[-1]["restarts"],
            "converged": all(st["converged"] for st in stats),
            "true_rel2_residual": res["rel2"],
            "solve_ms_device": float(np.median([st["ms"] for st in stats])),
            "set_mesh_ms": setup_ms,
        }

    main_prec = PS.MIXED if args.precision == "mixed" else PS.FP64
    other_prec = PS.FP64 if main_prec == PS.MIXED else PS.MIXED
    s, d_x, main = run_variant(main_prec, args.steps, args.warmup)
    _, _, other = run_variant(other_prec, max(3, args.steps // 4), 1)
    sz = s.sizes()
    ms_per_step = main["ms_per_step"]
    # the same solve through the host-buffer entry point (b in and x out over
    # PCIe every solve): reported beside the device-resident value, not as it
    nh = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(nh):
        s.solve(b, rel_tol=1e-10, max_restarts=100)
    host_rate = nh / (time.perf_counter() - t0)
    # the pressure step around the solve, device-resident (b = rhs - L p_old,
    # the solve from 0, p = p_old + dp - <dp>_h2), p_old = the last solution
    d_rhs = PS.DeviceArray.from_host(s, b)
    d_p = PS.DeviceArray(s, mesh.n)
    p_half = 0.5 * d_x.numpy()
    t_ps = []
    for _ in range(nh):
        d_p.upload(p_half)
        t0 = time.perf_counter()
        st_ps = s.pressure_step_device(d_rhs, d_p, rel_tol=1e-10, max_restarts=100)
        t_ps.append(time.perf_counter() - t0)
    pressure_step = {"ms": float(np.median(t_ps)) * 1e3, "iterations": st_ps["iterations"],
                     "converged": st_ps["converged"],
                     "p_old": "half the solution: the step solves for the other half"}

    # every iteration kernel timed alone (HIP events on the solver's stream);
    # the roofline object is the dominant one
    kern = []
    for k in range(5):
        ms_k = s.kernel_time(k, reps=100)
        byt = kernel_bytes(k, sz["n"], mesh.n_blocks, main_prec == PS.FP64)
        kern.append({"kernel": KERNELS[k], "ms_per_launch": ms_k, "algorithmic_bytes": byt,
                     "GB_per_s": byt / (ms_k * 1e-3) / 1e9})
    dom = max(range(5), key=lambda k: kern[k]["ms_per_launch"])
    # SpMV alone on the bench mesh: stencil kernel and plain CSR kernel
    d_y = PS.DeviceArray(s, mesh.n)
    ms_spmv = s.spmv_time(d_x, d_y, kind=0, reps=args.spmv_reps)
    ms_csr = s.spmv_time(d_x, d_y, kind=1, reps=args.spmv_reps)
    bytes_spmv = spmv_bytes_csr(sz["n"], sz["nnz"])
    achieved = bytes_spmv / (ms_spmv * 1e-3) / 1e9
    fp64 = main if main_prec == PS.FP64 else other
    mixed = main if main_prec == PS.MIXED else other

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == PS.FP64 else "f32/bf16 inner, f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": "3-level graded AMR Poisson solve (16x8 base blocks of 8x8 cells, disk "
                        "(0.5,0.5) r=0.1, radii scaled x5), seeded 16-vortex RHS + 8 disk sources, "
                        "FP64 BiCGSTAB + 64x64 block-Jacobi to 1e-10 rel 2-norm residual",
            "n_blocks": mesh.n_blocks, "n": sz["n"], "nnz": sz["nnz"],
            "explicit_rows": sz["n_explicit"],
            "precision": args.precision,
            "iterations": main["iterations"], "restarts": main["restarts"],
            "converged": main["converged"], "true_rel2_residual": main["true_rel2_residual"],
            "solve_ms_device": main["solve_ms_device"],
            "graph": args.graph,
            "fp64": fp64,
            "mixed": mixed,
            "mixed_speedup_vs_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
            "host_buffers_solves_per_s": host_rate,
            "pressure_step_device": pressure_step,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kern[dom]["kernel"] + (" (FP64)" if main_prec == PS.FP64 else " (mixed)"),
            "achieved": kern[dom]["GB_per_s"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kern[dom]["GB_per_s"] / HBM_PEAK_GBS,
            "traffic": pmc_traffic(f"iter{dom}") if main_prec == PS.FP64 else None,
            "algorithmic_bytes": kern[dom]["algorithmic_bytes"],
            "ms_per_launch": kern[dom]["ms_per_launch"],
            "profile": "profiles/roofline_kernels.md, profiles/pmc_traffic.txt",
        },
        "iteration_kernels": kern,
        "roofline_spmv": {
            "bound": "hbm", "kernel": "k_spmv (stencil SpMV, FP64, bench mesh)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": pmc_traffic("spmv"),
            "algorithmic_bytes_csr_equivalent": bytes_spmv, "ms_per_launch": ms_spmv,
            "csr_kernel_ms_per_launch": ms_csr,
        },
    }
    if not args.no_batch:
        out["config"]["amr6_1M"] = bench_amr6(PS, W, args.batch_steps, args.graph)
        batch, roof_dram = bench_batch(PS, W, args.batch_steps, args.graph)
        out["config"]["batch"] = batch
        out["roofline_dram"] = roof_dram
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, b)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
