"""AddressSanitizer + UndefinedBehaviorSanitizer builds on the CPU: the oracle
(oracle/tm_oracle.c, every front-end on small machines) and the host's layout
and size arithmetic (csrc/tm_layout.h, shared with the engine's host code),
each driven by a program under tests/sanitize/.  A sanitizer report or a
failed check fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")


def build_and_run(tmp_path, compiler, sources, extra, name):
    if shutil.which(compiler) is None:
        pytest.skip("%s is not installed" % compiler)
    exe = str(tmp_path / name)
    r = subprocess.run([compiler, *SAN, *extra, *sources, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("the sanitizer runtime is not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, (
        r.stdout[-2000:] + r.stderr[-4000:])
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    out = build_and_run(tmp_path, "gcc", [os.path.join(HERE, "sanitize", "oracle_driver.c")],
                        ["-std=gnu11", "-fopenmp", "-lm"], "oracle_driver")
    assert out.count("plain type") == 22 and out.count("individualized type") == 2, out


def test_host_layout_under_sanitizers(tmp_path):
    out = build_and_run(tmp_path, "g++", [os.path.join(HERE, "sanitize", "layout_driver.cpp")],
                        ["-std=c++17", "-I", os.path.join(ROOT, "lua-santoku-tsetlin_amd", "csrc")], "layout_driver")
    assert "0 failures" in out, out
