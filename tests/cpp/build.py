"""Builds the C++ facade test (tests/cpp/_build/facade_test) against
libsstg.so and the oracle library.  Called by __graft_entry__.build()."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "facade_test")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    lib_dir = os.path.join(ROOT, "sstgraph_amd")
    orc_dir = os.path.join(ROOT, "oracle", "_build")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'oracle')}",
           os.path.join(HERE, "facade_test.cpp"), "-o", OUT,
           f"-L{lib_dir}", "-lsstg", f"-L{orc_dir}", "-lsstg_oracle",
           "-Wl,-rpath,$ORIGIN/../../../sstgraph_amd",
           "-Wl,-rpath,$ORIGIN/../../../oracle/_build"]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
