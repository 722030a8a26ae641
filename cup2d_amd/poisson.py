"""Host-side interface of the solver, mirroring the reference's plugin
boundary for the pressure-Poisson path:

* ``PoissonSolver`` with ``solve(input, output)``      -- source/Poisson/Base.h:8-15
* ``makePoissonSolver(sim)`` string factory             -- source/Poisson/Base.cpp:8-30
* the Python binding ``poisson.Solver.solve(input=, output=)`` -- source/Bindings/Poisson.cpp:7-15

``HipPoissonSolver`` plays the role of ``ExpAMRSolver`` (the reference's
"cuda_iterative" solver, ExpAMRSolver.cpp:354-416): it reads the solver
parameters from the simulation data, re-assembles the operator when the mesh
changed (the reference keys this on ``pres->UpdateFluxCorrection``,
ExpAMRSolver.cpp:370-381), solves on the GPU through the C ABI and writes the
mean-free solution into ``output``.

``Solver`` is the thin object wrapper of the C ABI used by the tests and the
benchmark.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .mesh import BLEN, Mesh

FP64, MIXED = 0, 1
NORM_INF, NORM_L2 = 0, 1


def _check(h, rc):
    if rc != 0:
        msg = _lib.load().cup2d_ps_last_error(h).decode()
        raise _lib.SolverError(f"cup2d_ps error {rc}: {msg}")


def default_config(**kw) -> _lib.Config:
    cfg = _lib.Config()
    _lib.load().cup2d_ps_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise TypeError(f"unknown config field {k}")
        setattr(cfg, k, v)
    return cfg


def _mesh_desc(mesh: Mesh):
    keep = [np.ascontiguousarray(mesh.level, dtype=np.int32),
            np.ascontiguousarray(mesh.bx, dtype=np.int32),
            np.ascontiguousarray(mesh.by, dtype=np.int32)]
    d = _lib.MeshDesc(mesh.bpdx, mesh.bpdy, mesh.levels, int(mesh.periodic_x), int(mesh.periodic_y),
                      float(mesh.extent), mesh.n_blocks, keep[0].ctypes.data, keep[1].ctypes.data,
                      keep[2].ctypes.data)
    return d, keep


def assemble_csr(mesh: Mesh, mean_constraint=0):
    """Host-only assembly of the operator's FP64 CSR (cup2d_op_csr).  Returns
    (row_ptr int64, col int32, val float64, mean_row)."""
    lib = _lib.load()
    d, keep = _mesh_desc(mesh)
    mr = ctypes.c_longlong()
    nnz = lib.cup2d_op_csr(ctypes.byref(d), int(mean_constraint), None, None, None, 0, ctypes.byref(mr))
    if nnz == -1:
        raise _lib.SolverError(lib.cup2d_ps_last_error(None).decode())
    nnz = -nnz
    rp = np.empty(mesh.n + 1, np.int64)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    got = lib.cup2d_op_csr(ctypes.byref(d), int(mean_constraint), rp.ctypes.data, col.ctypes.data,
                           val.ctypes.data, nnz, ctypes.byref(mr))
    assert got == nnz
    return rp, col, val, mr.value


class DeviceArray:
    """Device buffer (FP64 by default) owned through a solver handle."""

    def __init__(self, solver: "Solver", n: int, dtype=np.float64):
        self.solver = solver
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(solver.h, _lib.load().cup2d_ps_dmalloc(solver.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, solver, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(solver, a.size, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(self.solver.h, _lib.load().cup2d_ps_h2d(self.solver.h, self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        _check(self.solver.h, _lib.load().cup2d_ps_d2h(self.solver.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            _lib.load().cup2d_ps_dfree(self.solver.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.solver.h:
                self.free()
        except Exception:
            pass


class Solver:
    """One solver handle (one HIP stream, its device buffers, one mesh)."""

    def __init__(self, config: Optional[_lib.Config] = None, **kw):
        lib = _lib.load()
        self.config = config if config is not None else default_config(**kw)
        h = ctypes.c_void_p()
        rc = lib.cup2d_ps_create(ctypes.byref(self.config), ctypes.byref(h))
        if rc != 0:
            raise _lib.SolverError(f"cup2d_ps_create error {rc}: {lib.cup2d_ps_last_error(None).decode()}")
        self.h = h.value
        self.mesh: Optional[Mesh] = None

    def close(self):
        if self.h:
            _lib.load().cup2d_ps_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- mesh ------------------------------------------------------------
    def set_mesh(self, mesh: Mesh, on_device: bool = False):
        """Assembles the operator of `mesh`: on the host (default, keeps a
        plain CSR copy) or by GPU kernels (on_device=True, same operator bit
        for bit, no CSR copy)."""
        d, self._keep = _mesh_desc(mesh)
        lib = _lib.load()
        fn = lib.cup2d_ps_set_mesh_on_device if on_device else lib.cup2d_ps_set_mesh
        _check(self.h, fn(self.h, ctypes.byref(d)))
        self.mesh = mesh
        return self

    def sizes(self):
        n, nnz, ne, mr = (ctypes.c_longlong() for _ in range(4))
        _check(self.h, _lib.load().cup2d_ps_sizes(self.h, ctypes.byref(n), ctypes.byref(nnz),
                                                   ctypes.byref(ne), ctypes.byref(mr)))
        return {"n": n.value, "nnz": nnz.value, "n_explicit": ne.value, "mean_row": mr.value}

    def csr(self):
        s = self.sizes()
        rp = np.empty(s["n"] + 1, np.int64)
        col = np.empty(s["nnz"], np.int32)
        val = np.empty(s["nnz"], np.float64)
        _check(self.h, _lib.load().cup2d_ps_get_csr(self.h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return rp, col, val

    # -- solve -----------------------------------------------------------
    def solve(self, b, x0=None, tol=0.0, rel_tol=1e-10, max_restarts=100):
        """Host arrays in, (x, stats) out; x is mean-free (ExpAMRSolver.cpp:383-415)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        st = _lib.Stats()
        _check(self.h, _lib.load().cup2d_ps_solve(self.h, b.ctypes.data, x.ctypes.data, float(tol),
                                                  float(rel_tol), int(max_restarts), ctypes.byref(st)))
        return x, st.as_dict()

    def solve_device(self, d_b: DeviceArray, d_x: DeviceArray, tol=0.0, rel_tol=1e-10, max_restarts=100):
        st = _lib.Stats()
        _check(self.h, _lib.load().cup2d_ps_solve_device(self.h, d_b.ptr, d_x.ptr, float(tol),
                                                         float(rel_tol), int(max_restarts),
                                                         ctypes.byref(st)))
        return st.as_dict()

    def pressure_step_device(self, d_rhs: DeviceArray, d_p: DeviceArray, tol=0.0, rel_tol=1e-10,
                             max_restarts=100):
        """The pressure step around the solve (PressureSingle.cpp:842-879) on
        device arrays: b = rhs - L p_old, dp solved from 0, p <- p_old + dp -
        <dp>_h2.  d_p holds p_old on entry and p on exit; d_rhs is kept."""
        st = _lib.Stats()
        _check(self.h, _lib.load().cup2d_ps_pressure_step_device(
            self.h, d_rhs.ptr, d_p.ptr, float(tol), float(rel_tol), int(max_restarts), ctypes.byref(st)))
        return st.as_dict()

    def pressure_rhs_device(self, d_u, d_v, d_ud, d_vd, d_chi, dt, d_rhs):
        _check(self.h, _lib.load().cup2d_ps_pressure_rhs_device(self.h, d_u.ptr, d_v.ptr, d_ud.ptr, d_vd.ptr,
                                                                d_chi.ptr, float(dt), d_rhs.ptr))

    def pressure_correct_device(self, d_p, dt, d_u, d_v):
        _check(self.h, _lib.load().cup2d_ps_pressure_correct_device(self.h, d_p.ptr, float(dt), d_u.ptr,
                                                                    d_v.ptr))

    def pressure_rhs(self, u, v, udef_u, udef_v, chi, dt):
        """The divergence RHS of the pressure equation (PressureSingle.cpp:324-349)
        for host arrays; cells on coarse-fine faces with the flux correction
        (PressureSingle.cpp:350-397)."""
        d = [DeviceArray.from_host(self, np.ascontiguousarray(a, dtype=np.float64))
             for a in (u, v, udef_u, udef_v, chi)]
        d_rhs = DeviceArray(self, d[0].n)
        _check(self.h, _lib.load().cup2d_ps_pressure_rhs_device(self.h, *[a.ptr for a in d], float(dt),
                                                                d_rhs.ptr))
        return d_rhs.numpy()

    def pressure_correct(self, p, u, v, dt):
        """The velocity correction (PressureSingle.cpp:122-210) for host arrays;
        returns the corrected (u, v)."""
        d_p = DeviceArray.from_host(self, np.ascontiguousarray(p, dtype=np.float64))
        d_u = DeviceArray.from_host(self, np.ascontiguousarray(u, dtype=np.float64))
        d_v = DeviceArray.from_host(self, np.ascontiguousarray(v, dtype=np.float64))
        _check(self.h, _lib.load().cup2d_ps_pressure_correct_device(self.h, d_p.ptr, float(dt), d_u.ptr,
                                                                    d_v.ptr))
        return d_u.numpy(), d_v.numpy()

    # ---- the rest of the projection time step (device arrays) ----------------

    def advect_diffuse_device(self, d_u, d_v, uinf, nu, dt):
        """advDiff::operator() (advDiff.cpp:238-282) in place on device arrays."""
        _check(self.h, _lib.load().cup2d_ps_advect_diffuse_device(self.h, d_u.ptr, d_v.ptr, float(uinf[0]),
                                                                  float(uinf[1]), float(nu), float(dt)))

    def penalize_device(self, d_chi, us, vs, lam, dt, d_u, d_v):
        """PressureSingle::penalize (PressureSingle.cpp:268-322), one rigid obstacle."""
        _check(self.h, _lib.load().cup2d_ps_penalize_device(self.h, d_chi.ptr, float(us), float(vs), float(lam),
                                                            float(dt), d_u.ptr, d_v.ptr))

    def chi_disk_device(self, cx, cy, r, d_chi):
        """The disk's characteristic function (PutObjectsOnGrid.cpp:70-96)."""
        _check(self.h, _lib.load().cup2d_ps_chi_disk_device(self.h, float(cx), float(cy), float(r), d_chi.ptr))

    def max_speed_device(self, d_u, d_v, uinf):
        """findMaxU (Helpers.cpp:155-207)."""
        out = ctypes.c_double()
        _check(self.h, _lib.load().cup2d_ps_max_speed_device(self.h, d_u.ptr, d_v.ptr, float(uinf[0]),
                                                             float(uinf[1]), ctypes.byref(out)))
        return out.value

    def block_score_device(self, d_u, d_v, d_chi, rtol):
        """Per-block refinement score (AdaptTheMesh.cpp:11-122): max |omega|,
        or 2 rtol at the obstacle's interface; a host array."""
        n_blocks = self.sizes()["n"] // 64
        out = np.empty(n_blocks, np.float64)
        _check(self.h, _lib.load().cup2d_ps_block_score_device(self.h, d_u.ptr, d_v.ptr, d_chi.ptr,
                                                               float(rtol), out.ctypes.data))
        return out

    def transfer_device(self, old_mesh, new_mesh, kind, src, d_old, d_new):
        """Fields d_old (on old_mesh) onto d_new (on new_mesh) by the plan of
        mesh.adapt."""
        (do, ko), (dn, kn) = _mesh_desc(old_mesh), _mesh_desc(new_mesh)
        descs = (do, dn)
        k = np.ascontiguousarray(kind, dtype=np.int32)
        sr = np.ascontiguousarray(src, dtype=np.int64)
        n = len(d_old)
        if (len(d_new) != n or k.shape != (new_mesh.n_blocks,) or sr.shape != (new_mesh.n_blocks, 4)
                or any(a.n != old_mesh.n for a in d_old) or any(a.n != new_mesh.n for a in d_new)):
            raise ValueError("transfer: fields or plan do not match the meshes")
        po = (ctypes.c_void_p * n)(*[a.ptr for a in d_old])
        pn = (ctypes.c_void_p * n)(*[a.ptr for a in d_new])
        _check(self.h, _lib.load().cup2d_ps_transfer_device(self.h, ctypes.byref(descs[0]),
                                                            ctypes.byref(descs[1]), k.ctypes.data,
                                                            sr.ctypes.data, n, po, pn))
        del ko, kn

    def advect_diffuse(self, u, v, uinf, nu, dt):
        """Host arrays in, (u, v) after the two-stage update out."""
        d_u = DeviceArray.from_host(self, np.ascontiguousarray(u, dtype=np.float64))
        d_v = DeviceArray.from_host(self, np.ascontiguousarray(v, dtype=np.float64))
        self.advect_diffuse_device(d_u, d_v, uinf, nu, dt)
        return d_u.numpy(), d_v.numpy()

    def pressure_step(self, rhs, p_old, tol=0.0, rel_tol=1e-10, max_restarts=100):
        """Host arrays in, (p, stats) out."""
        d_rhs = DeviceArray.from_host(self, np.ascontiguousarray(rhs, dtype=np.float64))
        d_p = DeviceArray.from_host(self, np.ascontiguousarray(p_old, dtype=np.float64))
        st = self.pressure_step_device(d_rhs, d_p, tol, rel_tol, max_restarts)
        return d_p.numpy(), st

    # -- operator pieces ---------------------------------------------------
    def spmv(self, x, kind=0):
        dx = DeviceArray.from_host(self, x)
        dy = DeviceArray(self, dx.n)
        _check(self.h, _lib.load().cup2d_ps_spmv_device(self.h, dx.ptr, dy.ptr, int(kind)))
        return dy.numpy()

    def spmv_time(self, d_x: DeviceArray, d_y: DeviceArray, kind=0, reps=50):
        ms = ctypes.c_double()
        _check(self.h, _lib.load().cup2d_ps_spmv_time(self.h, d_x.ptr, d_y.ptr, int(kind), int(reps),
                                                      ctypes.byref(ms)))
        return ms.value

    def kernel_time(self, which, reps=100):
        """ms per launch of iteration kernel `which` (0 p update + prec,
        1 nu = A z1, 2 s update + prec, 3 t = A z2, 4 x/r update)."""
        ms = ctypes.c_double()
        _check(self.h, _lib.load().cup2d_ps_kernel_time(self.h, int(which), int(reps), ctypes.byref(ms)))
        return ms.value

    def precond(self, r):
        dr = DeviceArray.from_host(self, r)
        dz = DeviceArray(self, dr.n)
        _check(self.h, _lib.load().cup2d_ps_precond_device(self.h, dr.ptr, dz.ptr))
        return dz.numpy()

    def precond_bf16(self, r):
        """BF16-MFMA block preconditioner of the mixed solver on FP32 data."""
        dr = DeviceArray.from_host(self, r, np.float32)
        dz = DeviceArray(self, dr.n, np.float32)
        _check(self.h, _lib.load().cup2d_ps_precond_bf16_device(self.h, dr.ptr, dz.ptr))
        return dz.numpy()

    def residual_device(self, d_b: DeviceArray, d_x: DeviceArray):
        out = (ctypes.c_double * 4)()
        _check(self.h, _lib.load().cup2d_ps_residual_device(self.h, d_b.ptr, d_x.ptr, out))
        r2, rinf, b2, binf = list(out)
        return {"r2": r2, "rinf": rinf, "b2": b2, "binf": binf,
                "rel2": r2 / b2 if b2 > 0 else 0.0, "relinf": rinf / binf if binf > 0 else 0.0}

    def residual(self, b, x):
        return self.residual_device(DeviceArray.from_host(self, b), DeviceArray.from_host(self, x))


def _check_b(h, rc):
    if rc != 0:
        msg = _lib.load().cup2d_psb_last_error(h).decode()
        raise _lib.SolverError(f"cup2d_psb error {rc}: {msg}")


class BatchSolver:
    """Batch handle (cup2d_psb_*): many independent meshes solved together.

    Vectors are the instances' vectors concatenated in instance order;
    ``offsets`` gives the first unknown of each instance.  Device buffers
    come from an internal single-solve handle on the same device."""

    def __init__(self, config: Optional[_lib.Config] = None, **kw):
        lib = _lib.load()
        self.config = config if config is not None else default_config(**kw)
        h = ctypes.c_void_p()
        rc = lib.cup2d_psb_create(ctypes.byref(self.config), ctypes.byref(h))
        if rc != 0:
            raise _lib.SolverError(f"cup2d_psb_create error {rc}: {lib.cup2d_psb_last_error(None).decode()}")
        self.h = h.value
        self.mem = Solver(default_config(device=self.config.device))
        self.meshes = []
        self.offsets = None

    def close(self):
        if self.h:
            _lib.load().cup2d_psb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_meshes(self, meshes, on_device: bool = False):
        """Assembles the instances' operators on host threads (default) or by
        GPU kernels in one pipeline (on_device=True, the same operators bit
        for bit)."""
        descs, keep = [], []
        for m in meshes:
            d, k = _mesh_desc(m)
            descs.append(d)
            keep.append(k)
        arr = (_lib.MeshDesc * len(descs))(*descs)
        lib = _lib.load()
        fn = lib.cup2d_psb_set_meshes_on_device if on_device else lib.cup2d_psb_set_meshes
        _check_b(self.h, fn(self.h, len(descs), arr))
        off = np.empty(len(descs) + 1, np.int64)
        _check_b(self.h, _lib.load().cup2d_psb_offsets(self.h, off.ctypes.data))
        self.meshes = list(meshes)
        self.offsets = off
        return self

    @property
    def n(self):
        return int(self.offsets[-1])

    def split(self, v):
        """Per-instance views of a concatenated vector."""
        return [v[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.meshes))]

    def device_array(self, a=None, n=None):
        if a is not None:
            return DeviceArray.from_host(self.mem, a)
        return DeviceArray(self.mem, self.n if n is None else n)

    def solve(self, b, x0=None, tol=0.0, rel_tol=1e-10, max_restarts=100):
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert b.size == self.n
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        st = (_lib.Stats * len(self.meshes))()
        _check_b(self.h, _lib.load().cup2d_psb_solve(self.h, b.ctypes.data, x.ctypes.data, float(tol),
                                                     float(rel_tol), int(max_restarts), st))
        return x, [s.as_dict() for s in st]

    def solve_device(self, d_b: DeviceArray, d_x: DeviceArray, tol=0.0, rel_tol=1e-10, max_restarts=100):
        st = (_lib.Stats * len(self.meshes))()
        _check_b(self.h, _lib.load().cup2d_psb_solve_device(self.h, d_b.ptr, d_x.ptr, float(tol),
                                                            float(rel_tol), int(max_restarts), st))
        return [s.as_dict() for s in st]

    def pressure_step_device(self, d_rhs: DeviceArray, d_p: DeviceArray, tol=0.0, rel_tol=1e-10,
                             max_restarts=100):
        """The pressure step around the solve for every instance (as
        Solver.pressure_step_device); d_p: p_old in, p out."""
        st = (_lib.Stats * len(self.meshes))()
        _check_b(self.h, _lib.load().cup2d_psb_pressure_step_device(
            self.h, d_rhs.ptr, d_p.ptr, float(tol), float(rel_tol), int(max_restarts), st))
        return [s.as_dict() for s in st]

    def pressure_rhs(self, u, v, udef_u, udef_v, chi, dt):
        """The divergence RHS for all instances (host arrays, concatenated)."""
        d = [self.device_array(np.ascontiguousarray(a, dtype=np.float64)) for a in (u, v, udef_u, udef_v, chi)]
        d_rhs = self.device_array()
        _check_b(self.h, _lib.load().cup2d_psb_pressure_rhs_device(self.h, *[a.ptr for a in d], float(dt),
                                                                   d_rhs.ptr))
        return d_rhs.numpy()

    def pressure_correct(self, p, u, v, dt):
        """The velocity correction for all instances; returns (u, v)."""
        d_p, d_u, d_v = (self.device_array(np.ascontiguousarray(a, dtype=np.float64)) for a in (p, u, v))
        _check_b(self.h, _lib.load().cup2d_psb_pressure_correct_device(self.h, d_p.ptr, float(dt), d_u.ptr,
                                                                       d_v.ptr))
        return d_u.numpy(), d_v.numpy()

    def autotune(self, d_b: DeviceArray):
        ib = ctypes.c_int()
        ms = ctypes.c_double()
        _check_b(self.h, _lib.load().cup2d_psb_autotune(self.h, d_b.ptr, ctypes.byref(ib), ctypes.byref(ms)))
        return ib.value, ms.value

    def item_blocks(self):
        return int(_lib.load().cup2d_psb_item_blocks(self.h))

    def xcd_run(self):
        """Items per XCD run of the operator kernels (1: plain order)."""
        return int(_lib.load().cup2d_psb_xcd_run(self.h))

    def sizes(self):
        n, nnz, ne = (ctypes.c_longlong() for _ in range(3))
        _check_b(self.h, _lib.load().cup2d_psb_sizes(self.h, ctypes.byref(n), ctypes.byref(nnz),
                                                     ctypes.byref(ne)))
        return {"n": n.value, "nnz": nnz.value, "n_explicit": ne.value}

    def spmv(self, x):
        dx = self.device_array(x)
        dy = self.device_array()
        _check_b(self.h, _lib.load().cup2d_psb_spmv_device(self.h, dx.ptr, dy.ptr))
        return dy.numpy()

    def spmv_time(self, d_x: DeviceArray, d_y: DeviceArray, reps=50):
        ms = ctypes.c_double()
        _check_b(self.h, _lib.load().cup2d_psb_spmv_time(self.h, d_x.ptr, d_y.ptr, int(reps),
                                                         ctypes.byref(ms)))
        return ms.value

    def kernel_time(self, which, reps=20):
        """ms per launch of batched iteration kernel `which` with every
        instance iterating (0 p update + prec, 1 nu = A z1, 2 s update + prec,
        3 t = A z2, 4 x/r update, 5 alpha, 6 omega, 7 state evaluation)."""
        ms = ctypes.c_double()
        _check_b(self.h, _lib.load().cup2d_psb_kernel_time(self.h, int(which), int(reps), ctypes.byref(ms)))
        return ms.value

    def copy_time(self, d_x: DeviceArray, d_y: DeviceArray, reps=50):
        """ms per launch of the library's streaming copy d_y = d_x over n."""
        ms = ctypes.c_double()
        _check_b(self.h, _lib.load().cup2d_psb_copy_time(self.h, d_x.ptr, d_y.ptr, int(reps),
                                                         ctypes.byref(ms)))
        return ms.value


# ---------------------------------------------------------------------------
# Mirror of the reference's solver plugin interface
# ---------------------------------------------------------------------------

@dataclass
class ScalarGrid:
    """Block grid of scalars: data[b, iy, ix] on the blocks of ``mesh``
    (the reference's ScalarGrid, Definitions.h:225-235).  ``UpdateFluxCorrection``
    is the mesh-changed flag the reference solver reads and clears
    (ExpAMRSolver.cpp:370-372)."""

    mesh: Mesh
    data: np.ndarray = None
    UpdateFluxCorrection: bool = True

    def __post_init__(self):
        if self.data is None:
            self.data = np.zeros((self.mesh.n_blocks, 8, 8), np.float64)


@dataclass
class SimulationData:
    """The solver parameters the reference reads from SimulationData
    (SimulationData.h:92-98, defaults Simulation.cpp:239-245)."""

    poissonSolver: str = "hip_iterative"
    PoissonTol: float = 1e-6
    PoissonTolRel: float = 0.0
    maxPoissonRestarts: int = 30
    maxPoissonIterations: int = 1000
    bMeanConstraint: int = 0
    step: int = 0
    verbose: bool = False
    device: int = 0
    tmp: Optional[ScalarGrid] = None
    pres: Optional[ScalarGrid] = None
    extra: dict = field(default_factory=dict)


class PoissonSolver:
    """Base.h:8-15."""

    def solve(self, input: ScalarGrid, output: ScalarGrid) -> None:  # noqa: A002
        raise NotImplementedError


class HipPoissonSolver(PoissonSolver):
    """GPU BiCGSTAB + 64x64 block-Jacobi behind PoissonSolver (the role of
    ExpAMRSolver, ExpAMRSolver.cpp:25-416)."""

    def __init__(self, sim: SimulationData, precision=FP64, norm=NORM_INF, **cfg):
        self.sim = sim
        self.solver = Solver(default_config(device=sim.device, precision=precision,
                                            mean_constraint=int(sim.bMeanConstraint != 0), norm=norm,
                                            max_iter=int(sim.maxPoissonIterations), **cfg))
        self.last_stats = None
        self._mesh = None          # the mesh whose operator the handle holds
        self.assemblies = 0        # operator assemblies so far

    def solve(self, input: ScalarGrid, output: ScalarGrid) -> None:  # noqa: A002
        sim = self.sim
        # tolerances of the first 10 steps (ExpAMRSolver.cpp:366-368)
        tol = 0.0 if sim.step < 10 else sim.PoissonTol
        rel = 0.0 if sim.step < 10 else sim.PoissonTolRel
        restarts = 100 if sim.step < 10 else sim.maxPoissonRestarts
        # the mesh is held by reference: an object identity test cannot be
        # fooled by an id() reused after garbage collection
        if output.UpdateFluxCorrection or self._mesh is not output.mesh:
            output.UpdateFluxCorrection = False
            # the operator is assembled by GPU kernels (the same matrix as the
            # host assembly, bit for bit); only the block list crosses PCIe
            self.solver.set_mesh(output.mesh, on_device=True)
            self._mesh = output.mesh
            self.assemblies += 1
        if input.data.shape != output.data.shape:
            raise ValueError("input and output grids differ")
        b = input.data.reshape(-1)
        x = output.data.reshape(-1)
        xs, st = self.solver.solve(b, x, tol, rel, restarts)
        output.data[...] = xs.reshape(output.data.shape)
        self.last_stats = st
        if sim.verbose:
            print(f"  [BiCGSTAB/HIP]: {st['iterations']} iterations, {st['restarts']} restarts, "
                  f"error {st['err_final']:.3e} / {st['err_init']:.3e}")


def makePoissonSolver(sim: SimulationData) -> PoissonSolver:
    """String factory of Base.cpp:8-30.  "hip_iterative" (and the reference's
    GPU name "cuda_iterative", for which this solver is the drop-in) give the
    HIP solver in FP64; "hip_mixed" the mixed-precision variant.  The CPU
    solver "iterative" is not part of this library."""
    name = sim.poissonSolver
    if name in ("hip_iterative", "cuda_iterative"):
        return HipPoissonSolver(sim, FP64)
    if name == "hip_mixed":
        return HipPoissonSolver(sim, MIXED)
    if name == "iterative":
        raise RuntimeError('Poisson solver: "iterative" (the CPU AMRSolver) is not provided by '
                           "cup2d_amd; use \"hip_iterative\"")
    raise ValueError(f'Poisson solver: "{name}" unrecognized!')


Solver_ = PoissonSolver  # name of the Python binding class (Bindings/Poisson.cpp:10)
