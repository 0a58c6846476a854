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


def _hdf5_link_args(tmp_path):
    """HDF5 from /opt/conda, linked through a directory holding only its
    library and zlib, so that the loader keeps the system libstdc++."""
    lib = "/opt/conda/lib/libhdf5.so.103"
    d = tmp_path / "hdf5lib"
    d.mkdir()
    for name in ("libhdf5.so.103", "libz.so.1"):
        os.symlink(os.path.join("/opt/conda/lib", name), str(d / name))
    return ["-I", "/opt/conda/include"], [lib, "-Wl,-rpath," + str(d)]


@pytest.mark.skipif(not os.path.exists("/opt/conda/include/hdf5.h"), reason="HDF5 not present")
def test_field_files_round_trip(tmp_path):
    """The on-disk format: <name>.stoch.h5 (dataset /stochastic, IEEE F64LE,
    dims slowest first) read back bit for bit in FP64 and FP32 and checked by
    h5dump, the XDMF description, the .trend coefficients and the .field
    configuration through report() and readINI."""
    inc, libs = _hdf5_link_args(tmp_path)
    exe = str(tmp_path / "io_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                           *inc, os.path.join(ROOT, "tests", "host", "io_check.cc"), "-o", exe,
                           *libs])
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr
    h5dump = "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        return
    head = subprocess.check_output([h5dump, "-H", str(tmp_path / "f3.stoch.h5")], text=True)
    assert 'DATASET "stochastic"' in head and "H5T_IEEE_F64LE" in head
    assert "SIMPLE { ( 3, 4, 5 ) / ( 3, 4, 5 ) }" in head
    body = subprocess.check_output([h5dump, "-d", "/stochastic", "-y", "-w", "0",
                                    str(tmp_path / "lin.stoch.h5")], text=True)
    data = body[body.index("DATA {") + 6: body.rindex("}")]
    vals = [float(v) for v in data.replace("\n", " ").replace("}", " ").split(",") if v.strip()]
    assert vals == [i + 0.25 for i in range(12)]
