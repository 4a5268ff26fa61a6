"""Builds libsstg.so (the HIP library behind include/sstg.h) for gfx950.

    python -m sstgraph_amd.build        # or __graft_entry__.build()

Every .hip file under csrc/ is compiled by hipcc, and every host-only .cpp
file by the host compiler, into _build/; the objects are linked into
sstgraph_amd/libsstg.so next to this file.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsstg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


# host-only sources (file readers over the public C ABI)
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread",
            f"-I{os.path.join(ROOT, 'include')}"]


def _compile(src):
    base, ext = os.path.splitext(os.path.basename(src))
    obj = os.path.join(BUILD, base + (".o" if ext == ".hip" else ".cpp.o"))
    if ext == ".hip":
        cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [CXX] + CXXFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))
    objs = []
    errors = []
    workers = min(len(srcs), max(1, (os.cpu_count() or 2) // 2), 8)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            if rc != 0:
                errors.append(f"--- {src}\n{out}")
            else:
                if verbose and out.strip():
                    print(out, file=sys.stderr)
                objs.append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
           "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
