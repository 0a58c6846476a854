"""Host-only parts of the C++ drop-in API (include/parafields_amd/*.hh),
compiled with g++ and run without a GPU."""
import os
import shutil
import subprocess

import numpy as np
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


def _build_batch_asan_check(tmp_path):
    """The library's host code (pf_api.hip, pf_fast_f32/f64.hip) compiled
    host-only with AddressSanitizer and linked with the stand-in runtime of
    tests/host/fake_hip.cc and the driver tests/host/batch_check.cc."""
    import re
    hipcc = "/opt/rocm/bin/hipcc"
    clang = "/opt/rocm/llvm/bin/clang++"
    csrc = os.path.join(ROOT, "parafields-core_amd", "csrc")
    inc = os.path.join(ROOT, "include")
    objs, procs = [], []
    for src in ("pf_api.hip", "pf_fast_f32.hip", "pf_fast_f64.hip"):
        obj = str(tmp_path / src.replace(".hip", ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen(
            [hipcc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fPIC",
             "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer",
             "-Wno-unused-command-line-argument", "-I", inc, "-c", os.path.join(csrc, src),
             "-o", obj], stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out = p.communicate()[0]
        assert p.returncode == 0, out.decode()[-3000:]
    # the host objects name the (absent) device code object: give it a definition
    syms = set()
    for obj in objs:
        nm = subprocess.run(["nm", "-u", obj], capture_output=True, text=True).stdout
        syms.update(re.findall(r"(__hip_fatbin\w*)", nm))
    stub = tmp_path / "fatbin_stub.cc"
    stub.write_text("".join(f'extern "C" {{ char {s}[64]; }}\n' for s in sorted(syms)))
    exe = str(tmp_path / "batch_check")
    subprocess.check_call(
        [clang, "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
         "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", inc,
         os.path.join(ROOT, "tests", "host", "fake_hip.cc"),
         os.path.join(ROOT, "tests", "host", "batch_check.cc"), str(stub), *objs, "-o", exe, "-lpthread"])
    return exe


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not available")
def test_batch_graph_and_solver_bookkeeping_under_asan(tmp_path):
    """No GPU: the 1000-field uneven batch of the benchmark (seed 3), one
    autotuned batch per size class, then CG, DCT/DST and 1-D DFT plans, all
    through the C ABI with AddressSanitizer on the library's host code and
    the stand-in runtime checking streams, events, captures and graphs."""
    import workloads as wl
    exe = _build_batch_asan_check(tmp_path)
    descs, shapes, _ = wl.uneven_batch(1000, seed=3)
    n = len(descs)
    (tmp_path / "descs.bin").write_bytes(b"".join(bytes(d) for d in descs))
    cls = np.array([max(s) for s in shapes], dtype=np.int32)
    (tmp_path / "classes.bin").write_bytes(cls.tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([exe, str(tmp_path / "descs.bin"), str(n), str(tmp_path / "classes.bin")],
                       capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]


def _write_png(path, pixels, channels):
    """8-bit PNG (grayscale or RGB) from rows of byte values, written with
    zlib: signature, IHDR, one IDAT with filter byte 0 per row, IEND."""
    import struct
    import zlib

    def chunk(kind, data):
        return (struct.pack(">I", len(data)) + kind + data +
                struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))

    h = len(pixels)
    w = len(pixels[0]) // channels
    color = {1: 0, 3: 2}[channels]
    raw = b"".join(b"\x00" + bytes(row) for row in pixels)
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0))
           + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)


def _png_link_args(tmp_path):
    """libpng from /opt/conda through a directory holding only it and zlib
    (conda's own libstdc++ stays off the search path)."""
    d = tmp_path / "pnglib"
    d.mkdir(exist_ok=True)
    for name in ("libpng16.so.16", "libz.so.1"):
        link = d / name
        if not link.exists():
            link.symlink_to(os.path.join("/opt/conda/lib", name))
    return (["-DPARAFIELDS_AMD_HAVE_PNG", "-I", "/opt/conda/include"],
            [str(d / "libpng16.so.16"), "-Wl,-rpath," + str(d)])


@pytest.mark.skipif(not os.path.exists("/opt/conda/include/png.h"), reason="libpng not present")
def test_image_trend_component(tmp_path):
    """Image trend component (trend.hh:679-836, pngreader.hh): at 48 points of
    a 2 x 1 domain the value is 0.5 + 2 (1 - p / 255) with p the pixel at row
    h - 1 - int(y / 1 * h), column int(x / 2 * w) (the red channel of an RGB
    image); generate draws its shift last and the .trend file keeps it."""
    rng = np.random.default_rng(4)
    h, w = 5, 7
    gray = rng.integers(0, 256, size=(h, w))
    rgb = rng.integers(0, 256, size=(h, w, 3))
    _write_png(str(tmp_path / "g.png"), [list(map(int, r)) for r in gray], 1)
    _write_png(str(tmp_path / "c.png"), [list(map(int, r.ravel())) for r in rgb], 3)
    inc, libs = _png_link_args(tmp_path)
    exe = str(tmp_path / "image_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), *inc,
                           os.path.join(ROOT, "tests", "host", "image_check.cc"), "-o", exe, *libs])
    r = subprocess.run([exe, str(tmp_path / "g.png"), str(tmp_path / "c.png"), str(tmp_path)],
                       capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr
    v = np.loadtxt(str(tmp_path / "values.txt"))
    for x, y, g, c in v:
        row, col = h - 1 - int(y / 1.0 * h), int(x / 2.0 * w)
        assert g == pytest.approx(0.5 + 2 * (1 - gray[row, col] / 255.), abs=1e-15)
        assert c == pytest.approx(0.5 + 2 * (1 - rgb[row, col, 0] / 255.), abs=1e-15)
