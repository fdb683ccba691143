"""The device BVH builder's kernels and host SAH (j3d_amd/csrc/bvh_gpu.hip)
compiled for the CPU and run serially under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/lbvh_cpu/check_lbvh.cpp): tree structure,
records and bounds on random, degenerate and NaN meshes, before any of it runs
on a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "j3d_amd", "csrc", "bvh_gpu.hip")
HERE = os.path.join(ROOT, "tests", "lbvh_cpu")


def _lbvh_body():
    s = open(SRC).read()
    a = s.index("namespace lbvh {")
    b = s.index("}  // namespace lbvh") + len("}  // namespace lbvh")
    return s[a:b] + "\n"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_device_builder_kernels_on_the_cpu(tmp_path):
    (tmp_path / "lbvh_body.inc").write_text(_lbvh_body())
    exe = tmp_path / "check_lbvh"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-ffp-contract=off", "-I" + str(tmp_path), "-I" + HERE, "-I" + os.path.join(ROOT, "j3d_amd", "csrc"),
           "-I" + os.path.join(ROOT, "include"), "-o", str(exe), os.path.join(HERE, "check_lbvh.cpp")]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:]
