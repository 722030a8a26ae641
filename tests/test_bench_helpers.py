"""The benchmark's byte counts and its reading of the committed traffic
counters (profiles/pmc_traffic.txt) on the CPU: the roofline figures it
prints are these numbers divided by measured times."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

N, NB = 13881344, 216896  # the 64-instance batch


def test_kernel_bytes():
    # FP64: p update 40, nu = A z1 24, s update 32, t = A z2 24, x / r update 72 bytes per cell
    assert [bench.kernel_bytes(k, N, NB) for k in range(5)] == [
        40 * N, 24 * N + 32 * NB, 32 * N, 24 * N + 32 * NB, 72 * N]
    # batched mixed: FP32 Krylov vectors, the pass's FP32 correction in the x update
    assert [bench.kernel_bytes(k, N, NB, False) for k in range(5)] == [
        24 * N, 12 * N + 32 * NB, 16 * N, 12 * N + 32 * NB, 36 * N]
    # single mixed solve: x and x_opt stay FP64
    assert bench.kernel_bytes(4, N, NB, False, batch=False) == 48 * N


def test_spmv_minimal_bytes():
    # x and y once, 32 B per block, per explicit row its index and 9 coded entries
    assert bench.spmv_bytes_min(N, NB, 297045) == 16 * N + 32 * NB + 297045 * (4 + 6 * 9)


def test_pmc_traffic_reads_the_committed_summary():
    tr = bench.pmc_traffic("biter4")
    assert tr is not None
    # within a per cent of the algorithmic bytes: the x / r update moves nothing else
    assert tr == pytest.approx(bench.kernel_bytes(4, N, NB), rel=0.01)
    assert bench.pmc_traffic("no_such_mode") is None
