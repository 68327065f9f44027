"""The engine's transform draws come from LazyMt19937 (mpsft_amd/csrc/mt.h),
which must give std::mt19937's stream word for word (src/base.cc:32-36):
compiled with g++ and compared over 8000 streams (tests/mt_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_lazy_mt19937_matches_std(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "mt_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe,
                        os.path.join(ROOT, "tests", "mt_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["mismatches", "0", "streams", "8000"], r.stdout
