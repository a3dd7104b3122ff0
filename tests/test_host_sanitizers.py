"""The library's host code (octree and lists, plan, translation and order
tables) built with AddressSanitizer and UndefinedBehaviourSanitizer and run
on uniform, clustered, spherical, planar and degenerate inputs
(tests/sanitize/host_check.cpp).  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_code_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", HERE, "run"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=900)
    print(r.stdout[-4000:])
    assert r.returncode == 0
    assert "host_check: 0 failed checks" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
