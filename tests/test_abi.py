"""The C-ABI library loads and exports every symbol include/sstg.h declares;
argument errors are reported without touching a GPU.  CPU only."""
import ctypes
import os
import re
import subprocess

import pytest

import sstgraph_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "sstg.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(sstg_\w+)\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 30
    L = ctypes.CDLL(S.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(sstg_\w+)\b", out))
    assert set(names) <= exported


def test_abi_version_and_strerror():
    assert S.lib().sstg_abi_version() == 2
    for code in [0, -1, -2, -3, -5, -6]:
        assert S.strerror(code)
    assert S.strerror(-2) == "vertex id out of range"


def test_null_handle_and_bad_arguments_are_errors():
    L = S.lib()
    assert L.sstg_destroy(None) == S.SSTG_E_INVAL
    assert L.sstg_insert_batch(None, None, None, 5, 0) == S.SSTG_E_INVAL
    assert L.sstg_num_edges(None, None) == S.SSTG_E_INVAL
    assert L.sstg_bfs(None, 0, None, None, 0) == S.SSTG_E_INVAL
    assert L.sstg_create(10, 10, 0, None) == S.SSTG_E_INVAL


def test_create_without_gpu_reports_no_device():
    if S.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(S.DeviceUnavailable):
        S.SparseMatrixV(100)
    with pytest.raises(S.DeviceUnavailable):
        S.rmat(1024, 0, 16)


def test_product_never_references_the_oracle():
    # the product path must not route through the CPU oracle
    pat = re.compile(r"import oracle|from oracle|sstg_oracle|orc_[a-z]")
    for dirpath, _, files in os.walk(os.path.join(ROOT, "sstgraph_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not pat.search(text), f
    for f in os.listdir(os.path.join(ROOT, "include")):
        p = os.path.join(ROOT, "include", f)
        if os.path.isfile(p):
            assert not pat.search(open(p).read()), f


def test_cpp_facade_builds_and_skips_without_gpu():
    from tests.cpp import build as cpp_build
    out = cpp_build.build()
    r = subprocess.run([out], capture_output=True, text=True, timeout=60)
    if S.device_count() == 0:
        assert r.returncode == 77 and "skipped" in r.stdout


def test_ids_outside_u32_are_refused_before_any_call():
    """Ids of a wider or signed integer type must fit in [0, 2^32): they are
    refused rather than wrapped onto a valid vertex (no GPU needed)."""
    import numpy as np
    with pytest.raises(S.IdOutOfRange):
        S._u32(np.array([1, -1], dtype=np.int64))
    with pytest.raises(S.IdOutOfRange):
        S._u32(np.array([1 << 32], dtype=np.int64))
    with pytest.raises(TypeError):
        S._u32(np.array([1.5]))
    ok = S._u32(np.array([0, (1 << 32) - 1], dtype=np.int64))
    assert ok.dtype == np.uint32 and ok.tolist() == [0, (1 << 32) - 1]
    assert S._u32(np.zeros(0, dtype=np.int64)).size == 0
