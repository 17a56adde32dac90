"""The C++ drop-in API (include/gsap/partition.hpp) compiled into a GSAP.cu-shaped driver."""
import os
import subprocess

import numpy as np
import pytest

import gsap_amd
from conftest import GOLDEN, ROOT, tsv_graph

TSV = os.path.join(GOLDEN, "graphchallenge", "LL", "1000_nodes.tsv")


def build_driver(tmp_path):
    exe = str(tmp_path / "dropin")
    lib_dir = os.path.join(ROOT, "gsap_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "include", "gsap"), os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"),
           "-L", lib_dir, "-lgsap_amd", f"-Wl,-rpath,{lib_dir}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_dropin_builds_and_fails_loudly_without_device(tmp_path):
    exe = build_driver(tmp_path)
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present; covered by the gpu test")
    r = subprocess.run([exe, TSV], capture_output=True, text=True)
    assert r.returncode == 1 and "gsap_create" in r.stderr


@pytest.mark.gpu
def test_dropin_matches_python_api(tmp_path):
    exe = build_driver(tmp_path)
    r = subprocess.run([exe, TSV], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = tsv_graph("LL")
    with gsap_amd.Context(0) as ctx:
        lab, _ = ctx.partition(g, gsap_amd.Options())
    h = 1469598103934665603
    for b in lab.astype(np.uint32).tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert f"labels {h:016x}" in r.stdout
