"""Host-side checks of the device arena's bookkeeping (rgf_amd/csrc/rgf_arena.hpp),
compiled with g++ against a host-memory backend: no overlapping or misaligned blocks,
free ranges coalesced, a held working set re-served without growing, cap trimming and
the out-of-memory retry."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_bookkeeping(tmp_path):
    exe = tmp_path / "arena_check"
    src = os.path.join(ROOT, "tests", "native", "arena_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-I", os.path.join(ROOT, "rgf_amd", "csrc"), src, "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("g++ failed:\n" + r.stderr)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "arena checks passed" in r.stdout
