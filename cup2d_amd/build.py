"""Builds the native libraries of the package in-tree:

* libcup2d_mesh.so    -- quadtree generator (host C++, g++)
* libcup2d_poisson.so -- the solver and the device assembly (HIP for gfx950 + host C++, hipcc)

Run as ``python -m cup2d_amd.build``; __graft_entry__.build() calls build().
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed-FP32 instructions are left out of the device code (-packed-fp32-ops).
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build_mesh():
    out = os.path.join(HERE, "libcup2d_mesh.so")
    _run(["g++", *CXX_FLAGS, "-shared", "-o", out, os.path.join(CSRC, "mesh.cpp")])
    return out


def build_poisson():
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in ("assembly.cpp",):
        obj = os.path.join(bdir, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in ("poisson.hip", "batch.hip", "assemble_device.hip"):
        obj = os.path.join(bdir, src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    out = os.path.join(HERE, "libcup2d_poisson.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build():
    build_mesh()
    build_poisson()


if __name__ == "__main__":
    build()
    print("built", os.path.join(HERE, "libcup2d_mesh.so"), os.path.join(HERE, "libcup2d_poisson.so"))
