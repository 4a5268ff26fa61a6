"""Helpers for the GPU parity tests."""
import sys

import numpy as np
import pytest

import sstgraph_amd as S


def need_gpu():
    if S.device_count() == 0:
        pytest.skip("no GPU visible")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same_csr(g, og, what=""):
    off, cols = g.export_csr()
    ooff, ocols = og.csr()
    assert off.dtype == np.uint64 and cols.dtype == np.uint32
    assert np.array_equal(off, ooff), f"{what}: row offsets differ"
    assert np.array_equal(cols, ocols), f"{what}: row contents differ"
    assert g.M() == og.num_edges()


def rel_close(got, want, rtol, atol_floor=0.0):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    err = np.abs(got - want)
    tol = rtol * np.abs(want) + atol_floor
    bad = err > tol
    return (not bad.any()), float((err / np.maximum(np.abs(want), 1e-300)).max())
