"""Block-quadtree meshes: the (level, bx, by) block description the solver
takes in place of Cubism's grid (reference source/Poisson/ExpAMRSolver.cpp:186-249
reads exactly these fields from cubism::BlockInfo).

Generation runs in ``libcup2d_mesh.so`` (cup2d_amd/csrc/mesh.cpp).
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

BS = 8  # cells per block side
BLEN = BS * BS

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libcup2d_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.cup2d_mesh_graded.restype = ctypes.c_longlong
        lib.cup2d_mesh_graded.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
        ]
        lib.cup2d_mesh_adapt.restype = ctypes.c_longlong
        lib.cup2d_mesh_adapt.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_longlong,
        ]
        lib.cup2d_mesh_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


@dataclass
class Mesh:
    """Leaf blocks of a 2:1-balanced block quadtree.

    ``level[b], bx[b], by[b]`` give block b's refinement level and its index
    at that level; block b owns unknowns ``64*b .. 64*b+63`` (cell
    ``iy*8+ix``), the row order of the reference (ExpAMRSolver.h:74-75).
    """

    bpdx: int
    bpdy: int
    levels: int
    extent: float
    level: np.ndarray
    bx: np.ndarray
    by: np.ndarray
    periodic_x: bool = False
    periodic_y: bool = False
    meta: dict = field(default_factory=dict)

    @property
    def n_blocks(self) -> int:
        return int(self.level.shape[0])

    @property
    def n(self) -> int:
        return self.n_blocks * BLEN

    def h(self) -> np.ndarray:
        """Cell size per block: extent / (max(bpdx,bpdy)*8) / 2^level."""
        h0 = self.extent / (max(self.bpdx, self.bpdy) * BS)
        return h0 / np.exp2(self.level.astype(np.float64))

    def h2(self) -> np.ndarray:
        return self.h() ** 2

    def cell_centers(self):
        """(x, y) of every unknown, shape (n,), in row order."""
        h = self.h()
        ix = np.tile(np.arange(BS), BS)
        iy = np.repeat(np.arange(BS), BS)
        x = (self.bx[:, None] * BS + ix[None, :] + 0.5) * h[:, None]
        y = (self.by[:, None] * BS + iy[None, :] + 0.5) * h[:, None]
        return x.reshape(-1), y.reshape(-1)

    def cell_h(self) -> np.ndarray:
        return np.repeat(self.h(), BLEN)

    def domain(self):
        s0 = self.extent / max(self.bpdx, self.bpdy)
        return self.bpdx * s0, self.bpdy * s0


def uniform(bpdx: int, bpdy: int, extent: float = 1.0, periodic_x=False, periodic_y=False) -> Mesh:
    return graded(bpdx, bpdy, 1, extent, [], [], periodic_x, periodic_y)


def graded(bpdx, bpdy, levels, extent, disks, bands, periodic_x=False, periodic_y=False) -> Mesh:
    """Quadtree refined around disks: a level-l block is split while its
    distance to a disk is below bands[l]; then 2:1 balanced, Hilbert ordered."""
    lib = _load()
    d = np.ascontiguousarray(np.asarray(disks, dtype=np.float64).reshape(-1))
    bnd = np.ascontiguousarray(np.zeros(max(levels - 1, 1)) if len(bands) == 0 else
                               np.asarray(bands, dtype=np.float64))
    if levels > 1 and bnd.shape[0] < levels - 1:
        raise ValueError("need levels-1 bands")
    nd = d.shape[0] // 3
    args = (bpdx, bpdy, levels, float(extent), int(periodic_x), int(periodic_y), nd,
            d.ctypes.data if nd else None, bnd.ctypes.data)
    nb = lib.cup2d_mesh_graded(*args, None, None, None, 0)
    if nb < 0:
        raise ValueError(lib.cup2d_mesh_last_error().decode())
    lev = np.empty(nb, np.int32)
    bx = np.empty(nb, np.int32)
    by = np.empty(nb, np.int32)
    nb2 = lib.cup2d_mesh_graded(*args, lev.ctypes.data, bx.ctypes.data, by.ctypes.data, nb)
    assert nb2 == nb
    return Mesh(bpdx, bpdy, levels, float(extent), lev, bx, by, bool(periodic_x), bool(periodic_y),
                meta={"disks": d.reshape(-1, 3).tolist(), "bands": bnd.tolist()})


def from_blocks(bpdx, bpdy, levels, extent, blocks, periodic_x=False, periodic_y=False) -> Mesh:
    """Mesh from an explicit list of (level, bx, by) triples (kept in order)."""
    a = np.asarray(blocks, dtype=np.int32).reshape(-1, 3)
    return Mesh(bpdx, bpdy, levels, float(extent), a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(),
                bool(periodic_x), bool(periodic_y))


def adapt(mesh: Mesh, tags):
    """The mesh adapted to per-block tags (+1 refine, -1 compress, 0 leave;
    cup2d_mesh_adapt in csrc/mesh.cpp).  Returns (new mesh, src_kind, src):
    per new block, 0 = the old block src[b, 0] unchanged, 1 = refined from
    the old ancestor src[b, 0], 2 = compressed from the four old children
    src[b, 0..3] (quadrant (x & 1) + 2 (y & 1))."""
    lib = _load()
    t = np.ascontiguousarray(np.asarray(tags, dtype=np.int8))
    if t.shape[0] != mesh.n_blocks:
        raise ValueError("one tag per block")
    lev, bx, by = (np.ascontiguousarray(a, dtype=np.int32) for a in (mesh.level, mesh.bx, mesh.by))
    args = (mesh.bpdx, mesh.bpdy, mesh.levels, int(mesh.periodic_x), int(mesh.periodic_y), mesh.n_blocks,
            lev.ctypes.data, bx.ctypes.data, by.ctypes.data, t.ctypes.data)
    nb = lib.cup2d_mesh_adapt(*args, None, None, None, None, None, 0)
    if nb < 0:
        raise ValueError(lib.cup2d_mesh_last_error().decode())
    nl, nx, ny = (np.empty(nb, np.int32) for _ in range(3))
    kind = np.empty(nb, np.int32)
    src = np.empty((nb, 4), np.int64)
    nb2 = lib.cup2d_mesh_adapt(*args, nl.ctypes.data, nx.ctypes.data, ny.ctypes.data, kind.ctypes.data,
                               src.ctypes.data, nb)
    if nb2 != nb:
        raise ValueError(lib.cup2d_mesh_last_error().decode())
    m = Mesh(mesh.bpdx, mesh.bpdy, mesh.levels, mesh.extent, nl, nx, ny, mesh.periodic_x, mesh.periodic_y)
    return m, kind, src
