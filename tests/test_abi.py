"""CPU checks of the product library: it loads, exports every symbol its
headers declare, and its host-side assembly equals the oracle's bit for bit."""

import ctypes
import os
import re

import numpy as np
import pytest

from cup2d_amd import _lib
from cup2d_amd import mesh as M
from cup2d_amd import poisson as PS
from cup2d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cup2d_\w+)\s*\(", txt)))


def test_poisson_lib_exports_header_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared("cup2d_poisson.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_mesh_lib_exports_header_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "cup2d_amd", "libcup2d_mesh.so"))
    for n in _declared("cup2d_mesh.h"):
        assert hasattr(lib, n), n


def test_create_without_gpu_fails_cleanly():
    """No device here: creation must fail with an error code, not crash or fall back."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from cup2d_amd import poisson as PS, _lib\n"
            "try:\n    PS.Solver()\nexcept _lib.SolverError as e:\n    print('ERR', e)\n"
            "else:\n    print('OK')\n") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    out = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    assert out.startswith("ERR") or out == "OK", r.stdout + r.stderr


@pytest.mark.parametrize("which", ["c1", "c2", "c3small", "periodic", "tiny"])
@pytest.mark.parametrize("mc", [0, 1])
def test_product_assembly_equals_oracle(oracle, which, mc):
    if which == "c1":
        m = W.uniform_16x8_mesh()
    elif which == "c2":
        m = W.graded_disk_mesh()
    elif which == "c3small":
        m = W.four_disk_amr6_mesh(target=2500)
    elif which == "periodic":
        m = M.graded(8, 4, 4, 2.0, [(0.1, 0.1, 0.05), (1.9, 0.9, 0.04)], [0.3, 0.15, 0.08], True, True)
    else:
        m = M.from_blocks(2, 2, 2, 1.0, [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 2), (1, 1, 2),
                                         (1, 0, 3), (1, 1, 3)])
    rp, col, val, mr = PS.assemble_csr(m, mc)
    A = oracle.assemble(m, mc)
    assert np.array_equal(rp, A.row_ptr)
    assert np.array_equal(col, A.col)
    assert np.array_equal(val, A.val)  # bit-identical coefficients
    assert mr == A.mean_row


def test_unbalanced_mesh_rejected():
    # a level-2 block next to a level-0 block: no valid coarse-fine row
    m = M.from_blocks(2, 1, 3, 2.0, [(0, 0, 0)] + [(2, 4 + a, b) for a in range(4) for b in range(4)])
    with pytest.raises(_lib.SolverError):
        PS.assemble_csr(m)


def test_factory_strings():
    sim = PS.SimulationData(poissonSolver="iterative")
    with pytest.raises(RuntimeError):
        PS.makePoissonSolver(sim)
    sim = PS.SimulationData(poissonSolver="bogus")
    with pytest.raises(ValueError):
        PS.makePoissonSolver(sim)


def test_batch_create_validates_config_before_device():
    """Invalid configurations are rejected with EINVAL before any device is
    touched; a valid one fails here with ENODEV (no GPU), never crashes."""
    import subprocess
    import sys
    code = ("import sys, ctypes; sys.path.insert(0, %r)\n"
            "from cup2d_amd import poisson as PS, _lib\n"
            "lib = _lib.load()\n"
            "def rc(**kw):\n"
            "    c = PS.default_config()\n"
            "    for k, v in kw.items():\n"
            "        if k == 'ib': c.item_blocks = v\n"
            "        else: setattr(c, k, v)\n"
            "    h = ctypes.c_void_p()\n"
            "    return lib.cup2d_psb_create(ctypes.byref(c), ctypes.byref(h))\n"
            "print(rc(max_iter=0), rc(inner_tol=-1.0), rc(inner_tol=1.0), rc(ib=257), rc(ib=-1),"
            " rc(precision=7), rc(norm=3), rc(ib=256), rc())\n") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    vals = [int(v) for v in r.stdout.split()]
    assert vals[:7] == [1] * 7, r.stdout + r.stderr
    assert all(v in (0, 5) for v in vals[7:]), vals  # OK on a GPU box, ENODEV here


def test_batch_null_handle_is_einval():
    lib = _lib.load()
    assert lib.cup2d_psb_solve(None, None, None, 0.0, 1e-10, 10, None) == 1
    assert lib.cup2d_psb_set_meshes(None, 1, None) == 1
    assert lib.cup2d_psb_offsets(None, None) == 1
    assert lib.cup2d_psb_destroy(None) == 0


def test_null_handles_are_einval_without_a_device():
    """Entry points reject a NULL handle with EINVAL before touching a device."""
    lib = _lib.load()
    assert lib.cup2d_ps_set_mesh_on_device(None, None) == 1
    assert lib.cup2d_ps_pressure_step_device(None, None, None, 0.0, 1e-10, 10, None) == 1
    assert lib.cup2d_ps_pressure_rhs_device(None, None, None, None, None, None, 1.0, None) == 1
    assert lib.cup2d_ps_pressure_correct_device(None, None, 1.0, None, None) == 1
    assert lib.cup2d_psb_set_meshes_on_device(None, 1, None) == 1
    assert lib.cup2d_psb_pressure_step_device(None, None, None, 0.0, 1e-10, 10, None) == 1


def test_batch_projection_null_handle_is_einval():
    lib = _lib.load()
    assert lib.cup2d_psb_pressure_rhs_device(None, None, None, None, None, None, 1.0, None) == 1
    assert lib.cup2d_psb_pressure_correct_device(None, None, 1.0, None, None) == 1


@pytest.mark.parametrize("lang", ["c99", "c++17"])
def test_header_compiles_and_links_from_c(lang, tmp_path):
    """A translation unit that includes include/cup2d_poisson.h, compiled as
    C and as C++ with warnings as errors, links every declared entry point
    against the library and runs its host-only calls (config defaults, the
    host assembly, error reporting) without a GPU."""
    import subprocess
    src = os.path.join(ROOT, "tests", "c_abi", "abi_check.c")
    exe = str(tmp_path / "abi_check")
    libdir = os.path.dirname(_lib.LIB_PATH)
    if lang == "c99":
        cmd = ["gcc", "-std=c99", "-x", "c"]
    else:
        cmd = ["g++", "-std=c++17", "-x", "c++"]
    cmd += ["-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"), src,
            "-L", libdir, "-lcup2d_poisson", "-Wl,-rpath," + libdir, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().startswith("ABI OK 592"), r.stdout + r.stderr
    # every entry point of the header is named in the C file
    names = set(re.findall(r"&(cup2d_\w+)", open(src).read()))
    assert names == set(_declared("cup2d_poisson.h"))
