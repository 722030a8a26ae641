"""Loader of the committed golden fixtures (tests/golden/)."""

import hashlib
import os

import numpy as np

from cup2d_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("tiny_cf", "uniform16x8", "graded3")


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    sh = d["shape"]
    m = M.from_blocks(int(sh[0]), int(sh[1]), int(sh[2]), float(d["extent"][0]),
                      np.stack([d["level"], d["bx"], d["by"]], 1), bool(sh[3]), bool(sh[4]))
    return m, {k: d[k] for k in d.files}


def digest(name):
    path = os.path.join(GOLDEN, name + ".csr.sha256")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return f.read().strip()


def csr_digest(row_ptr, col, val):
    h = hashlib.sha256()
    for a in (np.asarray(row_ptr, np.int64), np.asarray(col, np.int32), np.asarray(val, np.float64)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
