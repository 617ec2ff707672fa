"""The host worker pool behind the per-job host loops of lock-step batched proofs.

Builds tests/host_pool_check.cpp against p3-miden_amd/csrc/host_pool.hpp with g++
and runs it: every index once, concurrent callers, nested loops, exceptions.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_host_pool(tmp_path):
    exe = str(tmp_path / "host_pool_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-fsanitize=thread",
                    "-I", os.path.join(ROOT, "p3-miden_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host_pool_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host pool ok" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
