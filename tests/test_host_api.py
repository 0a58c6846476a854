"""Host-only parts of the C++ drop-in API (include/parafields_amd/*.hh),
compiled with g++ and run without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_trend_part(tmp_path):
    """TrendPart: component order, closed-form shapes, generate's draw order
    and scaling, variance scalings, generateUncorrelated's stream."""
    exe = str(tmp_path / "trend_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "host", "trend_check.cc"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_config_reader_periodization_keys(tmp_path):
    """make_desc: classical / merge / fold / cofold keys and their defaults."""
    exe = str(tmp_path / "config_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "host", "config_check.cc"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr
