"""The C++ drop-in facade (include/sstgraph_amd/SSTGraph.hpp) used the way
the reference harness uses SparseMatrixV<true> and EdgeMapVertexMap, checked
against the oracle inside tests/cpp/facade_test.cpp."""
import os
import subprocess

import pytest

from tests.cpp import build as cpp_build
from tests.gpu_common import need_gpu
from tests.graphs import GOLDEN

pytestmark = pytest.mark.gpu


def test_cpp_facade_matches_oracle():
    need_gpu()
    if not os.path.exists(cpp_build.OUT):
        cpp_build.build()
    r = subprocess.run([cpp_build.OUT, os.path.join(GOLDEN, "johnson8-2-4.mtx")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "facade ok" in r.stdout
    assert "there are 420 elements" in r.stdout
