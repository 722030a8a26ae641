"""Benchmark of the pressure-Poisson hot path on one MI355X.

One step = one batched solve of the largest workload that fits one GPU:
64 independent AMR Poisson instances (block counts log-uniform in
[512, 8192], seed 4; 216,896 blocks, 13,881,344 unknowns in all), each with
its own disk refinement and seeded vortex right-hand side, every instance
solved from x0 = 0 to 1e-10 relative residual (2-norm) by FP64 BiCGSTAB with
the 64x64 block-Jacobi preconditioner, b and x resident in HBM.  `value` is
FP64 solves per second over the timed steps.  Prints one JSON line.

Also measured in the same run (all in `config`):
* the same batch in mixed precision (FP32 Krylov vectors, BF16 MFMA
  preconditioner, FP64 residual replacement), and the like-for-like
  mixed / FP64 batched ratio;
* the same 64 solves one after another with the single-solve FP64 solver,
  fused (graph-replayed) and unfused (one operation per launch, the
  reference GPU loop's structure): the denominators of the batch speedups;
* every instance's FP64 true residual, and the error of the FP64 and the
  mixed solutions against the oracle's FP64 solve on a sample of instances;
* the single-mesh workloads (3-level graded mesh, 1448 blocks; 6-level
  mesh, 16,229 blocks) in FP64 and mixed precision;
* `roofline`: the dominant batched FP64 iteration kernel, timed live with
  HIP events on the solver's stream, with the counted HBM traffic of the
  committed rocprofv3 summary; the batched SpMV on its minimal bytes;
* `cpu_baseline`: the oracle's restatement of the reference CPU solver
  (pipelined BiCGSTAB with the block-Cholesky preconditioner,
  AMRSolver.cpp:149-490) on a bounded sample of the same instances, OpenMP
  on the host cores.
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
WORKLOAD = ("64 independent AMR Poisson instances (N_b log-uniform in [512, 8192], seed 4; 5 levels, "
            "one random disk each; seeded 16-vortex RHS), each solved from x0 = 0 to 1e-10 rel "
            "2-norm residual, BiCGSTAB + 64x64 block-Jacobi")

KERNELS = ["p update + block preconditioner", "nu = A z1 (+ dots)", "s update + block preconditioner",
           "t = A z2 (+ dots)", "x, r update (+ dots)"]
KERNEL_NAMES64 = ["kb_pupdate64", "kb_spmv_nu<double>", "kb_supdate64", "kb_spmv_t<double>",
                  "kb_xrupdate<double>"]


def kernel_bytes(which, n, n_blocks, fp64=True, batch=True):
    """Algorithmic HBM bytes of one launch of iteration kernel `which`: the
    vectors it must read and write once (s = 8 B FP64 / 4 B FP32 Krylov
    vectors; the x update moves the FP64 x and x_opt, 8 B each, except in
    the batched mixed solve, which updates the pass's FP32 correction and its
    best, 4 B each), plus 32 B of operator metadata per block for the
    operator kernels (explicit rows add < 1 %)."""
    s = 8 if fp64 else 4
    xs = 4 if (batch and not fp64) else 8
    per_cell = {0: 5 * s + (0 if fp64 else 4),   # r, p, nu in; p, z1 out (+ projected r)
                1: 3 * s, 2: 4 * s, 3: 3 * s,    # z, aux in; y out / r, nu in; r, z2 out
                4: 3 * xs + 6 * s}[which]        # x in/out, x_opt; z1, z2, r, t, rhat in;
                                                 # r out
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


def spmv_bytes_min(n, n_blocks, n_explicit, width=9):
    """Minimal bytes of the stencil-form SpMV: x and y once (16 B per row),
    32 B of block metadata (neighbour ids, explicit-row mask, h^2) per block,
    and per explicit (coarse-fine) row its index (4 B) and `width` coded
    entries (4 B column + 2 B coefficient code each; 7-9 entries is typical)."""
    return 16 * n + 32 * n_blocks + n_explicit * (4 + 6 * width)


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON result)."""
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


def host_cpu():
    """(model name, logical CPUs) of this host, from /proc/cpuinfo."""
    model, count = "unknown", 0
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                count += 1
    except OSError:
        pass
    return model, count


def c4_sample(meshes):
    """Instances of the CPU baseline and the error check: four of the batch,
    evenly spaced in instance order (their block counts span the range)."""
    return [0, 16, 32, 48]


def cpu_baseline(meshes, rhs, budget_s=30.0):
    """The oracle's restatement of the reference CPU solver (pipelined
    BiCGSTAB + block Cholesky, AMRSolver.cpp:149-490) on the sample
    instances, one solve each, OpenMP on every host thread, stopped early
    when the budget is spent."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    O.set_num_threads(threads)
    done, blocks, iters = [], 0, []
    t_total = 0.0
    for k in c4_sample(meshes):
        m, b = meshes[k], rhs[k]
        A = O.assemble(m)
        h2 = O.h2(m)
        t0 = time.perf_counter()
        _, st = O.pipelined_bicgstab(A, m, b, rel_tol=1e-10, max_iter=4000, max_restarts=100, h2v=h2)
        t_total += time.perf_counter() - t0
        done.append(k)
        blocks += m.n_blocks
        iters.append(st["iterations"])
        if t_total > budget_s:
            break
    model, _ = host_cpu()
    batch_blocks = sum(m.n_blocks for m in meshes)
    return {"value": len(done) / t_total, "unit": "solves/s", "cores": threads, "kind": "port",
            "cpu": model,
            "sample": f"instances {done} of the 64-instance batch ({blocks} blocks, "
                      f"{[meshes[k].n_blocks for k in done]} each), one solve each (pipelined "
                      f"BiCGSTAB + block Cholesky, AMRSolver.cpp:149-490, to 1e-10 rel 2-norm, "
                      f"{iters} iterations) in {t_total:.1f} s on {threads} OpenMP threads",
            "batch_equivalent_solves_per_s": len(meshes) / (t_total * batch_blocks / blocks)}


def oracle_errors(meshes, rhs, x_by_prec, offsets):
    """Relative L2 and max-norm error of the GPU solutions against the
    oracle's FP64 solve (plain BiCGSTAB to 1e-12, h^2-mean removed) on the
    sample instances, and the oracle's long-double true residuals."""
    from oracle import oracle as O
    out = {name: {"rel_l2_max": 0.0, "rel_max_max": 0.0, "true_rel2_max": 0.0} for name in x_by_prec}
    for k in c4_sample(meshes):
        m, b = meshes[k], rhs[k]
        A = O.assemble(m)
        h2 = O.h2(m)
        xr, _ = O.bicgstab(A, b, rel_tol=1e-12, norm="l2", max_iter=8000)
        xr = O.mean_remove(h2, xr)
        for name, x in x_by_prec.items():
            xk = x[offsets[k]:offsets[k + 1]]
            e = out[name]
            e["rel_l2_max"] = max(e["rel_l2_max"], float(np.linalg.norm(xk - xr) / np.linalg.norm(xr)))
            e["rel_max_max"] = max(e["rel_max_max"], float(np.abs(xk - xr).max() / np.abs(xr).max()))
            e["true_rel2_max"] = max(e["true_rel2_max"], O.true_residual(A, b, xk)["rel2"])
    for e in out.values():
        e["instances"] = c4_sample(meshes)
        e["oracle"] = "plain BiCGSTAB (BiCGSTAB.cu:412-620 restated) to 1e-12 rel 2-norm, mean removed"
    return out


def batch_spmv_handle(PS, meshes):
    """FP64 batch handle with 32 blocks per work item (the tuned size of the
    FP64 batch) for the operator kernel alone."""
    cfg = PS.default_config(norm=PS.NORM_L2)
    cfg.item_blocks = 32
    return PS.BatchSolver(cfg).set_meshes(meshes, on_device=True)


def batch_spmv_roofline(PS, meshes, reps):
    """The batched operator kernel alone (kb_spmv, FP64) on its minimal
    stencil-form bytes, with the counted traffic and the streaming-copy
    ceiling of the same size."""
    bs = batch_spmv_handle(PS, meshes)
    sz = bs.sizes()
    d_x = bs.device_array(np.random.default_rng(0).normal(size=bs.n))
    d_y = bs.device_array()
    ms = bs.spmv_time(d_x, d_y, reps=reps)
    ms_copy = bs.copy_time(d_x, d_y, reps=reps)
    nb = sum(m.n_blocks for m in meshes)
    byt = spmv_bytes_min(sz["n"], nb, sz["n_explicit"])
    ach = byt / (ms * 1e-3) / 1e9
    traffic = pmc_traffic("batch_spmv")
    return {"bound": "hbm", "kernel": "kb_spmv (batched stencil SpMV, FP64, 13.9M rows)",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "algorithmic_bytes": byt, "ms_per_launch": ms, "traffic": traffic,
            "traffic_GB_per_s": traffic / (ms * 1e-3) / 1e9 if traffic else None,
            "traffic_over_algorithmic": traffic / byt if traffic else None,
            "stream_copy_GB_per_s": 16 * sz["n"] / (ms_copy * 1e-3) / 1e9}


def timed_batch(PS, meshes, b, precision, steps, warmup, use_graph, profile=0):
    """Batched solves of the whole batch from x0 = 0: `warmup` untimed, then
    `steps` timed (device-resident b and x; wall clock around each solve,
    which synchronises on completion).  Returns (result dict, x, stats)."""
    bs = PS.BatchSolver(PS.default_config(precision=precision, norm=PS.NORM_L2, max_iter=4000,
                                          use_graph=use_graph, profile=profile))
    t0 = time.perf_counter()
    bs.set_meshes(meshes, on_device=True)  # device-side assembly of every instance
    setup_ms = (time.perf_counter() - t0) * 1e3
    d_b = bs.device_array(b)
    d_x = bs.device_array(np.zeros_like(b))
    ib, ms_it = bs.autotune(d_b)
    zeros = np.zeros_like(b)
    for _ in range(warmup):
        d_x.upload(zeros)
        bs.solve_device(d_b, d_x, rel_tol=1e-10)
    t = []
    for _ in range(steps):
        d_x.upload(zeros)
        t0 = time.perf_counter()
        st = bs.solve_device(d_b, d_x, rel_tol=1e-10)
        t.append(time.perf_counter() - t0)
    its = np.array([s["iterations"] for s in st], dtype=np.float64)
    nbk = np.array([m.n_blocks for m in meshes], dtype=np.float64)
    tot = float(np.sum(t))
    # every instance's FP64 true residual: r = b - A x with the batch's
    # operator (device SpMV), norms on the host
    x = d_x.numpy()
    r = b - bs.spmv(x)
    off = bs.offsets
    rel2 = [float(np.linalg.norm(r[off[k]:off[k + 1]]) / np.linalg.norm(b[off[k]:off[k + 1]]))
            for k in range(len(meshes))]
    relinf = [float(np.abs(r[off[k]:off[k + 1]]).max() / np.abs(b[off[k]:off[k + 1]]).max())
              for k in range(len(meshes))]
    # iterations weighted by instance size: an iteration of an instance costs
    # in proportion to its blocks, so this mean times the all-running
    # iteration time is the batch's work-proportional time
    w_its = float(np.dot(its, nbk) / nbk.sum())
    res = {"solves_per_s": len(meshes) * steps / tot, "ms_per_batch": tot * 1e3 / steps,
           "ms_per_batch_median": float(np.median(t)) * 1e3,
           "set_meshes_ms_device_assembly": setup_ms, "item_blocks": ib, "xcd_run": bs.xcd_run(),
           "ms_per_iteration_tuned": ms_it,
           "iterations_max": int(its.max()), "iterations_mean": float(its.mean()),
           "iterations_mean_weighted_by_blocks": w_its,
           "work_proportional_ms": w_its * ms_it,
           "refinements_max": int(max(s["refinements"] for s in st)),
           "all_converged": bool(all(s["converged"] for s in st)),
           "true_rel2_max": max(rel2), "true_relinf_max": max(relinf)}
    return res, x, st, bs


def sequential(PS, meshes, rhs, schedule, use_graph):
    """The batch's solves one after another with the single-solve FP64
    solver (schedule 0: fused, graph-replayed; 1: one operation per launch,
    the reference GPU loop's structure).  One warm-up solve per mesh."""
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000, use_graph=use_graph, schedule=schedule)
    tot, iters, conv = 0.0, [], True
    for m, bb in zip(meshes, rhs):
        s.set_mesh(m, on_device=True)
        d_b = PS.DeviceArray.from_host(s, bb)
        d_x = PS.DeviceArray(s, m.n)
        d_x.upload(np.zeros(m.n))
        s.solve_device(d_b, d_x, rel_tol=1e-10)  # warm-up (graph capture)
        d_x.upload(np.zeros(m.n))
        t0 = time.perf_counter()
        st = s.solve_device(d_b, d_x, rel_tol=1e-10)
        tot += time.perf_counter() - t0
        iters.append(st["iterations"])
        conv = conv and st["converged"]
    return {"solves_per_s": len(meshes) / tot, "ms_total": tot * 1e3,
            "iterations_mean": float(np.mean(iters)), "all_converged": bool(conv)}


def bench_single(PS, W, mesh, b, steps, use_graph):
    """FP64, mixed and unfused FP64 single solves of one mesh (device-resident
    b and x, from x0 = 0, median of `steps` after one warm-up)."""
    out = {"n_blocks": mesh.n_blocks, "n": mesh.n}
    xs = {}
    for name, prec, sched in (("fp64", PS.FP64, 0), ("mixed", PS.MIXED, 0), ("fp64_unfused", PS.FP64, 1)):
        s = PS.Solver(device=0, norm=PS.NORM_L2, max_iter=6000, use_graph=use_graph,
                      precision=prec, schedule=sched)
        t0 = time.perf_counter()
        s.set_mesh(mesh, on_device=True)
        setup_ms = (time.perf_counter() - t0) * 1e3
        d_b = PS.DeviceArray.from_host(s, b)
        d_x = PS.DeviceArray(s, mesh.n)
        zeros = np.zeros(mesh.n)
        d_x.upload(zeros)
        s.solve_device(d_b, d_x, rel_tol=1e-10)  # warm-up
        t = []
        for _ in range(steps if sched == 0 else max(1, steps // 3)):
            d_x.upload(zeros)
            t0 = time.perf_counter()
            st = s.solve_device(d_b, d_x, rel_tol=1e-10)
            t.append(time.perf_counter() - t0)
        sec = float(np.median(t))
        xs[name] = d_x.numpy()
        out[name] = {"solves_per_s": 1.0 / sec, "ms_per_solve": sec * 1e3, "set_mesh_ms": setup_ms,
                     "iterations": st["iterations"], "refinements": st["refinements"],
                     "restarts": st["restarts"], "converged": st["converged"],
                     "true_rel2_residual": st["true_rel2"]}
        if name == "fp64":
            out["fp64"]["kernel_ms"] = [s.kernel_time(k, reps=100) for k in range(5)]
            # the device-resident pressure step (b = rhs - L p_old, the solve
            # of the increment from 0, the h^2-mean-free update), p_old a
            # perturbed solution of this RHS
            x = d_x.numpy()
            p_old = 0.9 * x + 0.01 * np.abs(x).max() * np.sin(np.arange(mesh.n))
            d_p = PS.DeviceArray(s, mesh.n)
            t = []
            for _ in range(max(1, steps // 2) + 1):
                d_p.upload(p_old)
                t0 = time.perf_counter()
                sp = s.pressure_step_device(d_b, d_p, rel_tol=1e-10)
                t.append(time.perf_counter() - t0)
            out["fp64"]["pressure_step"] = {"ms": float(np.median(t[1:])) * 1e3,
                                            "iterations": sp["iterations"], "converged": sp["converged"]}
            del d_p
        del s, d_b, d_x
    out["mixed_speedup_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["fused_speedup_vs_unfused"] = out["fp64"]["solves_per_s"] / out["fp64_unfused"]["solves_per_s"]
    out["mixed_rel_l2_vs_fp64"] = float(np.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                        np.linalg.norm(xs["fp64"]))
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


def batch_kernels(PS, bs, n, n_blocks, fp64, reps=20):
    """Every batched iteration kernel alone (HIP events on the handle's
    stream, every instance iterating): ms per launch, algorithmic bytes and
    GB/s; with the counted traffic of the committed summary (FP64 rows
    biter0..4 of profiles/pmc_traffic.txt)."""
    out = []
    for k in range(5):
        ms = bs.kernel_time(k, reps=reps)
        byt = kernel_bytes(k, n, n_blocks, fp64)
        tr = pmc_traffic(f"biter{k}") if fp64 else pmc_traffic(f"miter{k}")
        out.append({"kernel": KERNELS[k], "ms_per_launch": ms, "algorithmic_bytes": byt,
                    "GB_per_s": byt / (ms * 1e-3) / 1e9, "traffic": tr})
    return out


def cylinder_flow(PS, W, steps, precision, cpu_steps=0, adapt_every=0):
    """The projection time step past the cylinder (r = 0.05 at (0.3, 0.5),
    U = 0.2, Re = 1000, CFL 0.4) from the 5-level mesh (~8.3k blocks),
    static or, with adapt_every > 0, adapted as AdaptTheMesh schedules it
    (every step of the first ten, then every adapt_every steps), every
    operator on the GPU: steps/s over `steps` steps after one untimed
    step from rest, and, with cpu_steps > 0, the relative
    L2 difference of velocity and pressure from the same first steps run on
    the CPU (oracle operators, FP64 BiCGSTAB to 1e-12)."""
    from cup2d_amd import flow as FLOW
    m = W.cylinder_mesh()
    flow = FLOW.CylinderFlow(m, precision=precision, rel_tol=1e-10)
    out = {"n_blocks": m.n_blocks, "n": m.n, "levels": int(m.levels)}
    if cpu_steps:
        from oracle import flow as FL
        from oracle import oracle as O
        from oracle import projection as PR
        A = O.assemble(m)
        w = np.repeat(O.h2(m), 64)
        chi = FL.chi_disk(m, *W.CYL_CENTER, W.CYL_RADIUS)
        u = np.zeros(m.n)
        v = np.zeros(m.n)
        p = np.zeros(m.n)
        z = np.zeros(m.n)
        cpu_s = 0.0
        for _ in range(cpu_steps):
            dt = FL.time_step(m, u, v, flow.uinf, flow.nu, flow.cfl)
            flow.step(dt)
            t0 = time.perf_counter()
            u, v = FL.advect_diffuse(m, u, v, flow.uinf, flow.nu, dt)
            u, v = FL.penalize(u, v, chi, flow.us, flow.vs, flow.lam, dt)
            b = PR.pressure_rhs(m, u, v, z, z, chi, dt) - O.spmv(A, p)
            dp, _ = O.bicgstab(A, b, rel_tol=1e-12, norm="l2", max_iter=8000)
            p = dp + (p - np.dot(w, dp) / np.sum(w))
            u, v = PR.pressure_correct(m, p, u, v, dt)
            cpu_s += time.perf_counter() - t0
        gu, gv, gp = flow.fields()
        out["cpu_compare"] = {
            "steps": cpu_steps,
            "cpu_ms_per_step": cpu_s * 1e3 / cpu_steps,
            "cpu_threads": O.num_threads(),
            "cpu_kind": "port (oracle flow and projection operators in numpy, FP64 BiCGSTAB to 1e-12 in C/OpenMP)",
            "rel_l2_u": float(np.linalg.norm(gu - u) / np.linalg.norm(u)),
            "rel_l2_v": float(np.linalg.norm(gv - v) / np.linalg.norm(v)),
            "rel_l2_p": float(np.linalg.norm(gp - p) / np.linalg.norm(p))}
    flow = FLOW.CylinderFlow(m, precision=precision, rel_tol=1e-10, adapt_every=adapt_every)
    flow.step()  # untimed first step (graph capture of the solver's iteration chunks)
    its, nbs = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = flow.step()
        its.append(st["iterations"])
        nbs.append(flow.mesh.n_blocks)
        if not st["converged"]:
            out["unconverged_step"] = flow.steps
    sec = time.perf_counter() - t0
    if adapt_every:
        out["adaptation"] = {"every": adapt_every, "adaptations": flow.adaptations,
                             "n_blocks_min": int(min(nbs)), "n_blocks_max": int(max(nbs)),
                             "n_blocks_final": int(nbs[-1]), "levels_final": int(flow.mesh.level.max()) + 1}
    gu, gv, _ = flow.fields()
    out.update({"steps": steps, "steps_per_s": steps / sec, "ms_per_step": sec * 1e3 / steps,
                "simulated_time": flow.time, "pressure_iterations_mean": float(np.mean(its)),
                "pressure_iterations_max": int(max(its)),
                "max_speed_final": float(max(np.abs(gu + flow.uinf[0]).max(), np.abs(gv).max()))})
    return out


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
    if which == "flow":
        # the flow operators of the time step on a uniform 128x128-block mesh
        # (1,048,576 cells, walls) and on the 5-level cylinder mesh (8348
        # blocks, coarse-fine faces).  Algorithmic bytes per cell: the
        # advection-diffusion step 160 (two stages, each: u, v in and the two
        # right-hand sides out (32), then u0, v0 and the right-hand sides in
        # and u, v out (48)); penalization 40 (chi, u, v in, u, v out); the
        # largest speed 16 (u, v in); the adaptation score 24 (u, v, chi in);
        # the transfer of one field to a refined copy of the mesh 16 per new
        # cell (8 in, 8 out)
        from cup2d_amd import mesh as M
        out = {}
        uni = M.uniform(128, 128, 1.0)
        uni.levels = 2  # room for the refined copy of the transfer
        for label, m in (("uniform_128x128", uni), ("cylinder_5_levels", W.cylinder_mesh())):
            s = PS.Solver().set_mesh(m, on_device=True)
            rng = np.random.default_rng(0)
            x, y = m.cell_centers()
            d_u = PS.DeviceArray.from_host(s, 0.1 * np.sin(6 * x) * np.cos(5 * y))
            d_v = PS.DeviceArray.from_host(s, -0.1 * np.cos(6 * x) * np.sin(5 * y))
            d_chi = PS.DeviceArray(s, m.n)
            s.chi_disk_device(0.3, 0.5, 0.05, d_chi)
            tags = np.ones(m.n_blocks, np.int8)
            m2, kind, src = M.adapt(m, tags)
            d_new = PS.DeviceArray(s, m2.n)
            res = {"cells": m.n}
            for name, fn, byt in (
                    ("advect_diffuse", lambda: s.advect_diffuse_device(d_u, d_v, (0.2, 0.0), 1e-4, 1e-4), 160 * m.n),
                    ("penalize", lambda: s.penalize_device(d_chi, -0.2, 0.0, 1e6, 1e-4, d_u, d_v), 40 * m.n),
                    ("max_speed", lambda: s.max_speed_device(d_u, d_v, (0.2, 0.0)), 16 * m.n),
                    ("block_score", lambda: s.block_score_device(d_u, d_v, d_chi, 5.0), 24 * m.n),
                    ("transfer_refine_all", lambda: s.transfer_device(m, m2, kind, src, [d_u], [d_new]), 16 * m2.n)):
                fn()
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                ms = (time.perf_counter() - t0) * 1e3 / reps
                res[name] = {"ms_per_call": ms, "algorithmic_bytes": byt, "GB_per_s": byt / (ms * 1e-3) / 1e9}
            out[label] = res
        print(json.dumps({"kernel": "flow", **out}))
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
        cfg.item_blocks = 32
        bs = PS.BatchSolver(cfg).set_meshes(meshes)
        d_b = bs.device_array(b)
        d_x = bs.device_array(np.zeros_like(b))
        st = bs.solve_device(d_b, d_x, rel_tol=1e-14)
        print(json.dumps({"kernel": which, "iterations": max(s["iterations"] for s in st),
                          "ms": st[0]["ms"]}))
        return
    if which[:5] in ("biter", "miter"):
        # batched iteration kernel K alone on the 64-instance batch (FP64:
        # biterK, mixed: miterK), every instance iterating, 32 blocks per item
        meshes = W.batch64_meshes()
        cfg = PS.default_config(precision=PS.FP64 if which[0] == "b" else PS.MIXED, norm=PS.NORM_L2)
        cfg.item_blocks = 32
        bs = PS.BatchSolver(cfg).set_meshes(meshes, on_device=True)
        k = int(which[5:])
        n, nb = bs.n, sum(m.n_blocks for m in meshes)
        ms = bs.kernel_time(k, reps=reps)
        byt = kernel_bytes(k, n, nb, which[0] == "b")
        print(json.dumps({"kernel": KERNELS[k], "ms_per_launch": ms, "launches": reps + 1,
                          "algorithmic_bytes": byt, "GB_per_s": byt / (ms * 1e-3) / 1e9}))
        return
    if which.startswith("iter") or which.startswith("xiter") or which.startswith("aiter"):
        # iteration kernel K of the 3-level mesh's single solve (iterK FP64,
        # xiterK mixed) or of the 6-level 16k-block mesh's (aiterK, FP64)
        mesh = W.four_disk_amr6_mesh() if which[0] == "a" else W.graded_disk_mesh()
        mixed = which[0] == "x"
        s = PS.Solver(norm=PS.NORM_L2, precision=PS.MIXED if mixed else PS.FP64)
        s.set_mesh(mesh)
        k = int(which[-1])
        ms = s.kernel_time(k, reps=reps)
        print(json.dumps({"kernel": KERNELS[k], "ms_per_launch": ms, "launches": reps + 2,
                          "algorithmic_bytes": kernel_bytes(k, mesh.n, mesh.n_blocks, not mixed, False)}))
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
    nb = sz["n"] // 64
    byt = spmv_bytes_min(sz["n"], nb, sz["n_explicit"])
    print(json.dumps({"kernel": which, "ms_per_launch": ms, "launches": reps, "n": sz["n"],
                      "nnz": sz["nnz"], "algorithmic_bytes": byt, "GB_per_s": byt / (ms * 1e-3) / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batched solves (each: 64 instances)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle-error legs")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--graph", type=int, default=1,
                    help="1: replay iteration chunks from a hipGraph; 0: plain stream launches")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the sequential solves and the single-mesh workloads")
    modes = ["spmv", "batch_spmv", "batch_fp64", "batch_mixed", "copy", "assembly", "projection", "flow"]
    modes += [f"{p}iter{k}" for p in ("", "x", "a", "b", "m") for k in range(5)]
    ap.add_argument("--kernel-only", choices=modes, default=None,
                    help="profiling aid: launch only this kernel --steps times and exit "
                         "(spmv: operator on the 3-level mesh; batch_spmv: batched operator on the "
                         "64-instance batch; batch_fp64 / batch_mixed: 64 iterations of the "
                         "batched solve, every kernel launched from the stream; iterK: iteration "
                         "kernel K of the 3-level mesh's solve (xiterK: mixed; aiterK: FP64 on the 6-level mesh); biterK / miterK: batched FP64 / "
                         "mixed iteration kernel K with every instance iterating; copy: the "
                         "library's streaming copy over 13,881,344 doubles, the bandwidth ceiling "
                         "and the traffic counters' calibration; assembly: host against device-side "
                         "assembly of the 6-level 16k-block mesh; projection / flow: the operators "
                         "either side of the pressure solve / the rest of the time step)")
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

    meshes = W.batch64_meshes()
    rhs = [W.workload_rhs(m, 400 + k) for k, m in enumerate(meshes)]
    b = np.concatenate(rhs)
    n_blocks = int(sum(m.n_blocks for m in meshes))
    progress(f"batch: {len(meshes)} instances, {n_blocks} blocks")
    fp64, x64, st64, bs64 = timed_batch(PS, meshes, b, PS.FP64, args.steps, args.warmup, args.graph)
    progress(f"batched FP64: {fp64['solves_per_s']:.1f} solves/s")
    offsets = bs64.offsets
    sz = bs64.sizes()
    kern = batch_kernels(PS, bs64, bs64.n, n_blocks, True)
    del bs64
    mixed, x32, _, bs32 = timed_batch(PS, meshes, b, PS.MIXED, args.steps, args.warmup, args.graph)
    progress(f"batched mixed: {mixed['solves_per_s']:.1f} solves/s")
    kern32 = batch_kernels(PS, bs32, bs32.n, n_blocks, False)
    del bs32
    cfg = {
        "workload": WORKLOAD, "instances": len(meshes), "n_blocks": n_blocks, "n": sz["n"],
        "nnz": sz["nnz"], "explicit_rows": sz["n_explicit"],
        "block_counts": {"min": int(min(m.n_blocks for m in meshes)),
                         "max": int(max(m.n_blocks for m in meshes))},
        "graph": args.graph, "fp64": fp64, "mixed": mixed,
        "mixed_batched_vs_fp64_batched": mixed["solves_per_s"] / fp64["solves_per_s"],
        "fp64_iteration_kernels": kern, "mixed_iteration_kernels": kern32,
    }
    if not args.no_extra:
        seq = sequential(PS, meshes, rhs, 0, args.graph)
        progress(f"sequential fused FP64: {seq['solves_per_s']:.1f} solves/s")
        unf = sequential(PS, meshes, rhs, 1, 0)
        progress(f"sequential unfused FP64: {unf['solves_per_s']:.1f} solves/s")
        cfg["fp64_sequential_fused"] = seq
        cfg["fp64_sequential_unfused"] = unf
        cfg["mixed_batched_vs_fp64_sequential_fused"] = mixed["solves_per_s"] / seq["solves_per_s"]
        cfg["mixed_batched_vs_fp64_sequential_unfused"] = mixed["solves_per_s"] / unf["solves_per_s"]
        cfg["fp64_batched_vs_fp64_sequential_fused"] = fp64["solves_per_s"] / seq["solves_per_s"]
        g = W.graded_disk_mesh()
        cfg["graded_1448_blocks"] = bench_single(PS, W, g, W.workload_rhs(g, 2, with_sources=True), 10,
                                                 args.graph)
        progress("graded 1448-block mesh done")
        a6 = W.four_disk_amr6_mesh()
        cfg["amr6_16k_blocks"] = bench_single(PS, W, a6, W.workload_rhs(a6, 3), 3, args.graph)
        progress("6-level 16k-block mesh done")
        cfg["assembly"] = bench_batch_assembly(PS, meshes)
        cfg["cylinder_flow_mixed"] = cylinder_flow(PS, W, 100, PS.MIXED, cpu_steps=0 if args.no_cpu else 2)
        progress("cylinder flow (mixed) done")
        cfg["cylinder_flow_fp64"] = cylinder_flow(PS, W, 100, PS.FP64)
        progress("cylinder flow (FP64) done")
        cfg["cylinder_flow_adaptive_mixed"] = cylinder_flow(PS, W, 100, PS.MIXED, adapt_every=20)
        progress("cylinder flow (adaptive, mixed) done")
    if not args.no_cpu:
        cfg["error_vs_oracle"] = oracle_errors(meshes, rhs, {"fp64": x64, "mixed": x32}, offsets)
        progress("errors against the oracle done")
    dom = max(range(5), key=lambda k: kern[k]["ms_per_launch"])
    out = {
        "metric": METRIC,
        "value": fp64["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL_NAMES64[dom] + ": " + kern[dom]["kernel"] + " (FP64, 64-instance batch)",
            "achieved": kern[dom]["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kern[dom]["GB_per_s"] / HBM_PEAK_GBS,
            "traffic": kern[dom]["traffic"],
            "algorithmic_bytes": kern[dom]["algorithmic_bytes"],
            "ms_per_launch": kern[dom]["ms_per_launch"],
            "profile": "profiles/roofline_batch.md, profiles/pmc_traffic.txt",
        },
        "roofline_spmv": batch_spmv_roofline(PS, meshes, args.spmv_reps),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(meshes, rhs)
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
