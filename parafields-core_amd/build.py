"""Build step of libparafields_amd.so (hipcc, gfx950) and of the host C++
test driver. Run as `python parafields-core_amd/build.py` or through
__graft_entry__.build()."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LIB = os.path.join(HERE, "lib", "libparafields_amd.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off (hipcc's host pass warns about the
    # unknown feature; that is harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build_library(verbose: bool = False) -> str:
    """Compile the three translation units in parallel and link them."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objdir = os.path.join(HERE, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = ["pf_api.hip", "pf_fast_f32.hip", "pf_fast_f64.hip"]
    procs, objs = [], []
    for src in srcs:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *[f for f in HIP_FLAGS if f != "-shared"], "-c", "-o", obj,
               os.path.join(HERE, "csrc", src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_host_driver(verbose: bool = False) -> str:
    """C++ host-API test driver (include/parafields_amd/*.hh over the C ABI)."""
    src = os.path.join(HERE, "host", "host_driver.cc")
    out = os.path.join(HERE, "lib", "pf_host_driver")
    if not os.path.exists(src):
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out,
           "-L", os.path.dirname(LIB), "-lparafields_amd", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build_oracle(verbose: bool = False) -> None:
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main():
    build_library(verbose=True)
    build_host_driver(verbose=True)
    build_oracle(verbose=True)


if __name__ == "__main__":
    sys.exit(main())
