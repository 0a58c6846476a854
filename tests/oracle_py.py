"""ctypes wrapper of the FP64 CPU oracle (oracle/pf_oracle.cc).

Test infrastructure: only tests, the smoke check and the benchmark's CPU
baseline use it, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libpforacle.so")
sys.path.insert(0, os.path.join(ROOT, "parafields-core_amd"))
import pf_binding as pfb  # noqa: E402

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = ctypes.CDLL(ORACLE_SO)
    vp, i32, i64, u64, dp = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                             ctypes.c_uint64, ctypes.c_double)
    D = ctypes.POINTER(pfb.PfDesc)
    sig = {
        "pfo_last_error": (ctypes.c_char_p, []),
        "pfo_set_threads": (i32, [i32]),
        "pfo_sizes": (i32, [D, ctypes.POINTER(pfb.PfSizes)]),
        "pfo_eigs": (i32, [D, vp, vp, ctypes.POINTER(pfb.PfEigStats)]),
        "pfo_generate": (i32, [D, vp, vp]),
        "pfo_matvec": (i32, [D, i32, vp, vp]),
        "pfo_multiply_inverse": (i32, [D, vp, vp, ctypes.POINTER(ctypes.c_int)]),
        "pfo_pcg": (i32, [D, i32, i32, dp, vp, vp, ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_double)]),
        "pfo_value_transform": (i32, [i32, dp, vp, ctypes.c_longlong]),
        "pfo_noise_cpp": (i32, [ctypes.c_uint, ctypes.c_longlong, vp]),
        "pfo_noise_philox": (i32, [u64, u64, u64, ctypes.c_longlong, vp]),
        "pfo_philox_raw": (i32, [vp, vp, ctypes.c_longlong, vp]),
        "pfo_cov_eval": (i32, [D, vp, ctypes.c_longlong, vp]),
        "pfo_bessel_k": (dp, [dp, dp]),
        "pfo_cov_grid": (i32, [D, vp]),
        "pfo_sigmoid": (dp, [i32, dp, dp, dp, i32]),
        "pfo_fold_integral": (dp, [i32, dp, dp, i32]),
        "pfo_taper_integral": (dp, [i32, dp, dp, dp, i32]),
        "pfo_pcg_learned": (i32, [D, vp, i32, i32, dp, vp, vp, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double)]),
        "pfo_spectral_targets": (i32, [D, i32, dp, vp]),
        "pfo_mlp": (dp, [vp, i32, vp, vp, vp, i32, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status):
    if status == 0:
        return
    msg = load().pfo_last_error().decode()
    if status == pfb.PF_ENEG:
        raise pfb.NegativeEigenvalueError(msg)
    raise RuntimeError(f"oracle error {status}: {msg}")


def set_threads(n: int):
    load().pfo_set_threads(n)


def sizes(desc) -> pfb.PfSizes:
    s = pfb.PfSizes()
    _check(load().pfo_sizes(ctypes.byref(desc), ctypes.byref(s)))
    return s


def eigs(desc, custom=None, allow_negative=False):
    """(eigenvalues in the reference storage layout, stats dict, status)."""
    s = sizes(desc)
    out = np.empty(int(s.eig_entries))
    st = pfb.PfEigStats()
    c = None if custom is None else np.ascontiguousarray(custom, dtype=np.float64)
    status = load().pfo_eigs(ctypes.byref(desc), None if c is None else c.ctypes.data,
                             out.ctypes.data, ctypes.byref(st))
    if status != 0 and not (allow_negative and status == pfb.PF_ENEG):
        _check(status)
    return out, st.as_dict(), status


def generate(desc, noise):
    """One realization from `noise`; for the DFT field one draw: the field
    and the spare field stacked (2, domain)."""
    s = sizes(desc)
    nz = np.ascontiguousarray(noise, dtype=np.float64)
    assert nz.size >= s.noise_per_field
    dft = desc.field_backend == pfb.FIELD_BACKENDS["dft"]
    out = np.empty(int(s.domain_size) * (2 if dft else 1))
    _check(load().pfo_generate(ctypes.byref(desc), nz.ctypes.data, out.ctypes.data))
    return out.reshape(2, -1) if dft else out


def matvec(desc, x, op="multiply"):
    xx = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(xx)
    _check(load().pfo_matvec(ctypes.byref(desc), pfb.OPS[op], xx.ctypes.data, out.ctypes.data))
    return out


def multiply_inverse(desc, b):
    bb = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(bb)
    it = ctypes.c_int()
    _check(load().pfo_multiply_inverse(ctypes.byref(desc), bb.ctypes.data, out.ctypes.data,
                                       ctypes.byref(it)))
    return out, it.value


def pcg(desc, b, mode="plain", maxit=1000, rel_tol=1e-8):
    bb = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(bb)
    it = ctypes.c_int()
    rr = ctypes.c_double()
    _check(load().pfo_pcg(ctypes.byref(desc), pfb.CG_MODES[mode], maxit, rel_tol, bb.ctypes.data,
                          out.ctypes.data, ctypes.byref(it), ctypes.byref(rr)))
    return out, it.value, rr.value


def value_transform(v, kind, param=0.0):
    vv = np.array(v, dtype=np.float64, copy=True)
    _check(load().pfo_value_transform(pfb.TRANSFORMS[kind], param, vv.ctypes.data, vv.size))
    return vv


def noise_cpp(seed: int, n: int):
    out = np.empty(n)
    load().pfo_noise_cpp(seed, n, out.ctypes.data)
    return out


def noise_philox(seed: int, realization: int, first: int, n: int):
    out = np.empty(n)
    load().pfo_noise_philox(seed, realization, first, n, out.ctypes.data)
    return out


def philox_raw(ctr, key, nblocks=1):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4 * nblocks, dtype=np.uint32)
    load().pfo_philox_raw(c.ctypes.data, k.ctypes.data, nblocks, out.ctypes.data)
    return out


def cov_eval(desc, pts):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    out = np.empty(p.shape[0])
    _check(load().pfo_cov_eval(ctypes.byref(desc), p.ctypes.data, p.shape[0], out.ctypes.data))
    return out


def cov_grid(desc):
    """Embedding grid before the forward transform (classical or merge),
    natural layout [C2][C1][C0]: the DCT half grid or the full extended
    grid (R2C matrix)."""
    dim = desc.dim
    N = [desc.cells[i] * (1 if desc.periodic else desc.embedding_factor) for i in range(dim)]
    C = [n // 2 + 1 for n in N] if desc.matrix_backend == pfb.MATRIX_BACKENDS["dct"] else N
    out = np.empty(int(np.prod(C)))
    _check(load().pfo_cov_grid(ctypes.byref(desc), out.ctypes.data))
    return out.reshape(C[::-1])


def sigmoid(kind, one, zero, x, rec=0):
    """kind: 'clamp', 'smooth' or 'smoothstep' (covariance.hh:147-290)."""
    k = {"clamp": 0, "smooth": 1, "smoothstep": 2}[kind]
    return load().pfo_sigmoid(k, one, zero, x, rec)


def fold_integral(sigmoid, beta, norm, rec=0):
    """Integral of sigmoid(1, beta, x) over [1, norm] (fold embedding);
    sigmoid: 'smoothstep' or 'smooth'."""
    k = {"smoothstep": 0, "smooth": 1}[sigmoid]
    return load().pfo_fold_integral(k, beta, norm, rec)


def taper_integral(sigmoid, one, zero, hi, rec=0):
    """Integral of sigmoid(one, zero, x) over [one, hi]."""
    k = {"smoothstep": 0, "smooth": 1}[sigmoid]
    return load().pfo_taper_integral(k, one, zero, hi, rec)


def bessel_k(nu, x):
    return load().pfo_bessel_k(nu, x)


def pcg_learned(desc, b, log_mult, maxit=1000, rel_tol=1e-8):
    bb = np.ascontiguousarray(b, dtype=np.float64)
    lm = np.ascontiguousarray(log_mult, dtype=np.float64)
    out = np.empty_like(bb)
    it = ctypes.c_int()
    rr = ctypes.c_double()
    _check(load().pfo_pcg_learned(ctypes.byref(desc), lm.ctypes.data, lm.size, maxit, rel_tol,
                                  bb.ctypes.data, out.ctypes.data, ctypes.byref(it),
                                  ctypes.byref(rr)))
    return out, it.value, rr.value


def spectral_targets(desc, nbins, floor_value):
    out = np.empty(nbins)
    _check(load().pfo_spectral_targets(ctypes.byref(desc), nbins, floor_value, out.ctypes.data))
    return out


def mlp(dims, params, x, y, grads=True):
    """(loss, grads or None, outputs) of the FP64 reference MLP."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    p = np.ascontiguousarray(params, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    yy = np.ascontiguousarray(y, dtype=np.float64)
    g = np.empty(p.size) if grads else None
    out = np.empty((xx.shape[0], dims[-1]))
    loss = load().pfo_mlp(d.ctypes.data, d.size, p.ctypes.data, xx.ctypes.data, yy.ctypes.data,
                          xx.shape[0], None if g is None else g.ctypes.data, out.ctypes.data)
    return loss, g, out
