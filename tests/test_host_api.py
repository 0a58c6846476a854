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


def _build_gpumatrix_check(tmp_path):
    import pf_binding as pfb
    exe = str(tmp_path / "gpumatrix_check")
    libdir = os.path.dirname(pfb.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "host", "gpumatrix_check.cc"), "-o", exe,
                           "-L", libdir, "-lparafields_amd", "-Wl,-rpath," + libdir])
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_gpumatrix_adapter_compiles_against_reference_shaped_traits(tmp_path):
    """include/parafields_amd/gpumatrix.hh as the IsoMatrix / AnisoMatrix of
    stand-ins shaped like the reference's RandomFieldTraits and StochasticPart
    (private members, friend declarations, backend is_same checks): every
    member instantiates and links against the library."""
    exe = _build_gpumatrix_check(tmp_path)
    r = subprocess.run([exe, "compile"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr
