"""The allocation-free solve (cup2d_ps_solve_device_noalloc), called through
the C ABI: the same solution, iteration count and convergence as
cup2d_ps_solve_device on the same handle, bit for bit (the same kernels in
the same order on a fixed work space), on two calls in a row -- the second
on a smaller mesh, so the shared work array holds the first call's larger
contents -- and the size, precision and alias limits rejected with an
error."""

import ctypes

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_N = 131072


def _noalloc(s, d_b, d_x, rel_tol=1e-10):
    st = _lib.Stats()
    rc = _lib.load().cup2d_ps_solve_device_noalloc(s.h, d_b.ptr, d_x.ptr, 0.0, float(rel_tol), 100,
                                                  ctypes.byref(st))
    return rc, st.as_dict()


def _pair(s, m, seed):
    """Both entry points from x0 = 0 on the handle's current mesh."""
    b = W.vortex_rhs(m, seed)
    d_b = PS.DeviceArray.from_host(s, b)
    d_x1 = PS.DeviceArray.from_host(s, np.zeros(m.n))
    d_x2 = PS.DeviceArray.from_host(s, np.zeros(m.n))
    st1 = s.solve_device(d_b, d_x1, rel_tol=1e-10)
    rc, st2 = _noalloc(s, d_b, d_x2)
    assert rc == 0, _lib.load().cup2d_ps_last_error(s.h)
    return d_x1.numpy(), st1, d_x2.numpy(), st2


def test_noalloc_equals_solve_device_twice():
    big = W.graded_disk_mesh()              # 1448 blocks, 92,672 unknowns
    small = W.uniform_16x8_mesh()           # 128 blocks
    assert small.n < big.n <= MAX_N
    s = PS.Solver(norm=PS.NORM_L2, max_iter=4000)
    for m, seed in ((big, 3), (small, 4)):
        s.set_mesh(m)
        x1, st1, x2, st2 = _pair(s, m, seed)
        assert st1["converged"] and st2["converged"]
        assert st1["iterations"] == st2["iterations"] and st1["restarts"] == st2["restarts"]
        assert np.array_equal(x1, x2), m.n_blocks
        assert st2["true_rel2"] <= 1.05e-10
        assert abs(st2["true_rel2"] - st1["true_rel2"]) <= 1e-6 * st1["true_rel2"]
    # the handle's own buffers are untouched by the allocation-free call
    x3, st3, _, _ = _pair(s, small, 4)
    assert np.array_equal(x3, x1) and st3["iterations"] == st1["iterations"]


def test_noalloc_limits():
    lib = _lib.load()
    # more unknowns than the work space holds (2049 blocks)
    m = M.uniform(64, 33, 2.0)
    assert m.n > MAX_N
    s = PS.Solver().set_mesh(m)
    d = PS.DeviceArray.from_host(s, np.zeros(m.n))
    d2 = PS.DeviceArray.from_host(s, np.zeros(m.n))
    rc, _ = _noalloc(s, d, d2)
    assert rc == 1  # CUP2D_PS_EINVAL
    assert b"at most 131072" in lib.cup2d_ps_last_error(s.h)
    # mixed precision, aliasing, no mesh
    sm = PS.Solver(precision=PS.MIXED).set_mesh(W.uniform_16x8_mesh())
    d = PS.DeviceArray.from_host(sm, np.zeros(8192))
    d2 = PS.DeviceArray.from_host(sm, np.zeros(8192))
    assert _noalloc(sm, d, d2)[0] == 1
    s2 = PS.Solver().set_mesh(W.uniform_16x8_mesh())
    d = PS.DeviceArray.from_host(s2, np.zeros(8192))
    d2 = PS.DeviceArray.from_host(s2, np.zeros(8192))
    assert _noalloc(s2, d, d)[0] == 1
    s3 = PS.Solver()
    assert lib.cup2d_ps_solve_device_noalloc(s3.h, d.ptr, d2.ptr, 0.0, 1e-10, 100, None) == 2  # ENOMESH
