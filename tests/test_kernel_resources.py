"""Register budgets of the hot train tiles, read from the compiler's resource
remarks (hipcc -Rpass-analysis=kernel-resource-usage on the device pass).

A CU of gfx950 holds four 512-thread workgroups of the train kernel only at
64 VGPRs or fewer and below about 80 SGPRs (a recorded residency measurement:
four fit at 70 SGPRs, three at 86); three at more.  The workload-2 tile
(train_kernel_r4<4, step queue>) is built to that budget, and the encoder
tile (64 lanes x 8 words) to 128 VGPRs for two workgroups per CU; neither
may spill.  A compiler or source change that breaks
a budget keeps results bit-exact but costs a workgroup per CU, so it is caught
here rather than in a benchmark.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lua-santoku-tsetlin_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def resources(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc is not installed")
    out = tmp_path_factory.mktemp("res") / "k.o"
    cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           "-Wno-unused-command-line-argument", "-x", "hip", "-c", "--offload-device-only",
           os.path.join(CSRC, "tm_kernels.hip"), "-o", str(out), "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    res, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            res[cur] = {}
            continue
        m = re.search(r"remark: +(TotalSGPRs|VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and cur is not None:
            res[cur].setdefault(m.group(1).split()[0], int(m.group(2)))
    return res


def pick(res, *parts):
    hits = [k for k in res if all(p in k for p in parts)]
    assert len(hits) == 1, (parts, hits)
    return res[hits[0]]


def test_workload2_tile_fits_four_workgroups_per_cu(resources):
    r = pick(resources, "train_kernel_r4ILi4ELb1E")
    assert r["VGPRs"] <= 64, r
    assert r["TotalSGPRs"] <= 78, r
    assert r["ScratchSize"] == 0, r


@pytest.mark.parametrize("mangled", ["train_kernelILi8ELi64ELi7ELb1ELb0ELb1E",    # encoder: FULL, scan
                                     "train_kernelILi8ELi64ELi7ELb1ELb1ELb1E",    # FULL, step queue
                                     "train_kernelILi4ELi64ELi7ELb1ELb0ELb1E"])   # FULL, scan
def test_full_wave_tiles_do_not_spill(resources, mangled):
    r = pick(resources, mangled)
    assert r["VGPRs"] <= 128 and r["ScratchSize"] == 0, r


def test_predict_kernel_has_no_scratch(resources):
    r = pick(resources, "predict_kernel")
    assert r["ScratchSize"] == 0, r
