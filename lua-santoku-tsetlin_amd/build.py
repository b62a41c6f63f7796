"""Build libtsetlin_amd.so in place (hipcc, gfx950 only).

    python lua-santoku-tsetlin_amd/build.py

Objects go to lua-santoku-tsetlin_amd/_build/, the library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libtsetlin_amd.so")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed-FP32 instruction selection anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-command-line-argument",
]
# (source, extra defines, object name); tm_train_states.hip is built once per
# plane count other than 7 (state bits 2..16 except 8)
SOURCES = [("tm_kernels.hip", [], "tm_kernels.o"), ("tm_api.cpp", [], "tm_api.o"), ("tm_mlp.hip", [], "tm_mlp.o"),
           ("tm_data.hip", [], "tm_data.o")]
SOURCES += [("tm_train_states.hip", ["-DTM_PLANES=%d" % m], "tm_train_planes_%d.o" % m)
            for m in range(1, 16) if m != 7]


def _compile(job):
    src, defs, name = job
    obj = os.path.join(BUILD, name)
    cmd = [HIPCC, *COMMON, *defs, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    workers = max(1, min(len(SOURCES), os.cpu_count() or 1, 16))
    # the longest compile first
    with ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
