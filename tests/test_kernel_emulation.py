"""Host builds of the generation kernels (no GPU needed).

`kf_gen_last` (parafields-core_amd/csrc/pf_fast.hpp) is compiled with g++
against a small HIP shim (tests/emu/hip/hip_runtime.h): every workgroup runs
as blockDim.x std::threads with std::barrier for __syncthreads, and each
buffer ends at a PROT_NONE guard page. The harness (tests/emu/gen_pass_check.cc)
compares the fused Philox pass (counter-based normals drawn inside the kernel,
main pass plus odd-row tail pass) against the same kernel fed the identical
normals from a precomputed array, over both storage orders, 2D and 3D, R2C and
DCT layouts and odd tile counts. An out-of-bounds access stops the run naming
the buffer and offset; a mismatch or a missed entry fails the comparison.

`tests/emu/dct_pass_check.cc` does the same for the three specialised
DCT/DST passes (pf_dctdst_fast.hpp, 64 KB tiles): every block of small
shapes, sampled blocks of the 1024^3 shape at its launch geometry (three
realizations per launch), and the in-kernel Philox normals against the same
passes fed those normals from an array.

Three device-only constructs are replaced by plain C++ with the same result
before compiling: the three-input XOR instruction (a ^ b ^ k), the empty asm
barrier that hides threadIdx from the optimiser, and the non-temporal store
(an ordinary store of the same two values).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parafields-core_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "emu")

SUBST = [
    (r'asm\("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"\(r\) : "v"\(a\), "v"\(b\), "s"\(k\)\);',
     "r = a ^ b ^ k;"),
    (r'asm volatile\("" : "\+v"\(t\)\);', ""),
    (r"using V = T __attribute__\(\(ext_vector_type\(2\)\)\);", "struct V { T x, y; };"),
    (r"__builtin_nontemporal_store\(w, reinterpret_cast<V\*>\(p\)\);",
     "*reinterpret_cast<V*>(p) = w;"),
    (r'#include "\.\./\.\./include/parafields_amd\.h"', '#include "../include/parafields_amd.h"'),
]


def _host_sources(dst):
    src = os.path.join(dst, "src")
    os.makedirs(src)
    os.makedirs(os.path.join(dst, "include"))
    shutil.copy(os.path.join(ROOT, "include", "parafields_amd.h"), os.path.join(dst, "include"))
    for name in os.listdir(CSRC):
        if not name.endswith(".hpp"):
            continue
        text = open(os.path.join(CSRC, name)).read()
        for pat, rep in SUBST:
            text = re.sub(pat, rep, text)
        code = "\n".join(l.split("//")[0] for l in text.splitlines())
        assert "asm" not in re.sub(r"\w*asm\w+|\w+asm\w*", "", code), f"unhandled inline asm in {name}"
        open(os.path.join(src, name), "w").write(text)
    return src


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_gen_pass_host_build(tmp_path):
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "gen_pass_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "gen_pass_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout
    assert r.stdout.count(" OK") >= 20


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_dct_passes_host_build(tmp_path):
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "dct_pass_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "dct_pass_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_dct_generic_passes_host_build(tmp_path):
    """The generic (runtime-size) DCT/DST passes of pf_dctdst.hpp at the tile
    the library gives them under an 80 KB LDS budget (8 lines of 1024, 73,792
    bytes), the geometry of the two 1024^3 runs that faulted on the GPU: every
    block of small shapes and sampled blocks of the 1024^3 shapes stay inside
    their buffers."""
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "dct_generic_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "dct_generic_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_dft_passes_host_build(tmp_path):
    """The DFT field passes of pf_dft.hpp (1-D one tile and four-step, 2D and
    3D in both storage orders, FP64 and FP32, odd first fields) and the 1-D
    product passes, every block, against direct O(N^2) sums."""
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "dft_pass_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "dft_pass_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout and r.stdout.count(" OK") >= 11


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_fused_product_pass_host_build(tmp_path):
    """The fused last-axis pass of the covariance product (kf_fused_mult),
    every block of 3D (both storage orders) and 2D shapes, FP64 and FP32,
    multiply / root / inverse-floor, against direct O(N^2) sums."""
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "fused_pass_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "fused_pass_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_even_part_eigenvalue_y_pass_host_build(tmp_path):
    """The real y pass of the 2D R2C eigenvalues (kf_eig_c2r_y): from the x
    half spectrum of rows 0 .. N1/2 of a point-symmetric array, every block
    of two realizations (own thresholds), FP32 and FP64 roots, with and
    without the FP64 eigenvalues, N1 from 32 to 2048, against direct 2D DFTs;
    classification counts summed over the block partials match exactly."""
    src = _host_sources(str(tmp_path))
    exe = str(tmp_path / "eig_y_check")
    cmd = ["g++", "-std=c++20", "-O2", "-pthread", "-w", "-I", EMU, "-I", src,
           os.path.join(EMU, "eig_y_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all OK" in r.stdout
    assert r.stdout.count(" OK") >= 21
