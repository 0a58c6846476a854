"""ctypes binding of the C ABI declared in include/parafields_amd.h.

This is the Python-side caller of libparafields_amd.so (the HIP library). It
does no arithmetic of its own: every operation is a call into the library,
and loading fails loudly when the library has not been built.

The configuration mirrors the reference's ParameterTree keys
(include/parafields/fieldtraits.hh:169-209, matrix.hh:223-318); `Desc`
defaults follow the reference defaults (embedding.factor 2,
embedding.threshold 1e-14, randomField.cgIterations 100, fftw.transposed
default).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libparafields_amd.so")

PF_OK, PF_ENEG, PF_EINVAL, PF_EHIP, PF_ENOMEM, PF_ESTATE = 0, -1, -2, -3, -4, -5

COVARIANCES = {
    "exponential": 0, "gaussian": 1, "spherical": 2, "separableExponential": 3,
    "matern": 4, "matern32": 5, "matern52": 6, "dampedOscillation": 7,
    "gammaExponential": 8, "cauchy": 9, "generalizedCauchy": 10, "cubic": 11,
    "whiteNoise": 12, "custom-iso": 13,
}
ANISOTROPY = {"none": 0, "axiparallel": 1, "geometric": 2}
MATRIX_BACKENDS = {"dct": 0, "r2c": 1, "dft": 2}
FIELD_BACKENDS = {"r2c": 0, "dctdst": 1, "dft": 2}
PRECISION = {"fp64": 0, "fp32": 1}
PERIODIZATION = {"classical": 0, "merge": 1, "fold": 2, "cofold": 3}
SIGMOID = {"smoothstep": 0, "smooth": 1}
SIGMOID_COMBINE = {"radialPlus": 0, "radial": 1, "tensor": 2}
TRANSFORMS = {"none": 0, "logNormal": 1, "foldedNormal": 2, "sign": 3, "boxCox": 4}
OPS = {"multiply": 0, "root": 1, "inverse_embed": 2}
CG_MODES = {"reference": 0, "plain": 1, "embed_pcg": 2, "learned_pcg": 3}
MEM_HOST, MEM_DEVICE = 0, 1


class NegativeEigenvalueError(RuntimeError):
    """Counterpart of parafields::NegativeEigenvalueError (exceptions.hh:8-15)."""


class PfDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("cells", ctypes.c_int32 * 3),
        ("extensions", ctypes.c_double * 3),
        ("embedding_factor", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
        ("covariance", ctypes.c_int32),
        ("variance", ctypes.c_double),
        ("anisotropy", ctypes.c_int32),
        ("corr_length", ctypes.c_double * 9),
        ("matern_nu", ctypes.c_double),
        ("exp_gamma", ctypes.c_double),
        ("cauchy_alpha", ctypes.c_double),
        ("cauchy_beta", ctypes.c_double),
        ("matrix_backend", ctypes.c_int32),
        ("field_backend", ctypes.c_int32),
        ("transposed", ctypes.c_int32),
        ("approximate", ctypes.c_int32),
        ("threshold", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("cg_iterations", ctypes.c_int32),
        ("nugget", ctypes.c_double),
        ("r2c_check_all", ctypes.c_int32),
        ("periodization", ctypes.c_int32),
        ("sigmoid", ctypes.c_int32),
        ("sigmoid_combine", ctypes.c_int32),
        ("merge_recursions", ctypes.c_int32),
        ("sigmoid_start", ctypes.c_double),
        ("sigmoid_end", ctypes.c_double),
        ("fold_max_factor", ctypes.c_double),
    ]


class PfEigStats(ctypes.Structure):
    _fields_ = [
        ("small", ctypes.c_int64),
        ("zero", ctypes.c_int64),
        ("negative", ctypes.c_int64),
        ("smallest", ctypes.c_double),
        ("entries", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PfSizes(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("cells", ctypes.c_int32 * 3),
        ("extended", ctypes.c_int32 * 3),
        ("transposed", ctypes.c_int32),
        ("domain_size", ctypes.c_int64),
        ("extended_size", ctypes.c_int64),
        ("eig_entries", ctypes.c_int64),
        ("noise_per_field", ctypes.c_int64),
    ]


class PfCgOpts(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("rel_tol", ctypes.c_double),
        ("bin_log_multipliers", ctypes.POINTER(ctypes.c_float)),
        ("nbins", ctypes.c_int32),
    ]


class PfTuning(ctypes.Structure):
    _fields_ = [
        ("tile_div", ctypes.c_int32),
        ("chunks", ctypes.c_int32),
        ("overlap", ctypes.c_int32),
        ("lines", ctypes.c_int32),
        ("radix_order", ctypes.c_int32),
        ("sub_batch", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PfTuneResult(ctypes.Structure):
    _fields_ = [
        ("best", PfTuning),
        ("ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("candidates", ctypes.c_int32),
        ("from_cache", ctypes.c_int32),
    ]


class PfCgStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("rel_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


def make_desc(dim: int, cells: Sequence[int], extensions: Optional[Sequence[float]] = None,
              covariance: str = "exponential", variance: float = 1.0,
              corr_length: Sequence[float] | float = 0.05, anisotropy: str = "none",
              factor: int = 2, periodic: bool = False, matrix: Optional[str] = None,
              field: Optional[str] = None, transposed: int = -1, approximate: bool = False,
              threshold: float = 1e-14, precision: str = "fp64", cg_iterations: int = 100,
              nugget: float = 0.0, matern_nu: float = 1.0, exp_gamma: float = 1.0,
              cauchy_alpha: float = 1.0, cauchy_beta: float = 1.0,
              r2c_check_all: bool = False, periodization: str = "classical",
              sigmoid: str = "smoothstep", sigmoid_combine: str = "radialPlus",
              merge_recursions: int = 0, sigmoid_start: float = 0.0,
              sigmoid_end: float = 0.0, fold_max_factor: float = 0.0) -> PfDesc:
    """Descriptor with the reference's defaults. `matrix=None` picks what
    RandomField would (randomfield.hh:153-164): DCT (DefaultIsoMatrix) for
    none/axiparallel, R2C (DefaultAnisoMatrix) for geometric anisotropy, and
    the DFT matrix and field for 1-D (matrix.hh:1660-1736); `field=None` is
    the R2C field (DFT in 1-D).
    periodization / sigmoid / sigmoid_combine take the reference's names
    (embedding.periodization, embedding.sigmoid, stochastic.sigmoidCombine);
    merge_recursions (embedding.mergeRecursions, foldRecursions or
    cofoldRecursions), sigmoid_start / sigmoid_end and fold_max_factor
    (foldMaxFactor or cofoldMaxFactor) 0 mean the reference defaults,
    merge_recursions -1 means zero recursions."""
    d = PfDesc()
    d.dim = dim
    for i in range(3):
        d.cells[i] = int(cells[i]) if i < dim else 1
        d.extensions[i] = float(extensions[i]) if (extensions is not None and i < dim) else 1.0
    d.embedding_factor = factor
    d.periodic = int(periodic)
    d.covariance = COVARIANCES[covariance]
    d.variance = variance
    d.anisotropy = ANISOTROPY[anisotropy]
    cl = [corr_length] if np.isscalar(corr_length) else list(corr_length)
    for i, v in enumerate(cl):
        d.corr_length[i] = float(v)
    d.matern_nu, d.exp_gamma = matern_nu, exp_gamma
    d.cauchy_alpha, d.cauchy_beta = cauchy_alpha, cauchy_beta
    if matrix is None:
        # 1-D: DFTMatrixBackend (matrix.hh:1660-1736)
        matrix = "dft" if dim == 1 else ("r2c" if anisotropy == "geometric" else "dct")
    if field is None:
        field = "dft" if dim == 1 else "r2c"
    d.matrix_backend = MATRIX_BACKENDS[matrix]
    d.field_backend = FIELD_BACKENDS[field]
    d.transposed = transposed
    d.approximate = int(approximate)
    d.threshold = threshold
    d.precision = PRECISION[precision]
    d.cg_iterations = cg_iterations
    d.nugget = nugget
    d.r2c_check_all = int(r2c_check_all)
    d.periodization = PERIODIZATION[periodization]
    d.sigmoid = SIGMOID[sigmoid]
    d.sigmoid_combine = SIGMOID_COMBINE[sigmoid_combine]
    d.merge_recursions = merge_recursions
    d.sigmoid_start, d.sigmoid_end = sigmoid_start, sigmoid_end
    d.fold_max_factor = fold_max_factor
    return d


_lib = None


def load(path: str = LIB_PATH):
    """Load libparafields_amd.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    # torch ships its own HIP runtime (SONAME libamdhip64.so.7, loaded by the
    # unversioned name). Loaded first, it also serves this library's
    # dependency; loaded after this library pulled in /opt/rocm's copy, it
    # becomes a second runtime in the process and finds no GPU. So torch, when
    # present, is imported before the library is opened.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    vp, i32, i64, u64, dp = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                             ctypes.c_double)
    sig = {
        "pf_plan_create": (i32, [ctypes.POINTER(PfDesc), ctypes.POINTER(vp)]),
        "pf_plan_destroy": (i32, [vp]),
        "pf_plan_sizes": (i32, [vp, ctypes.POINTER(PfSizes)]),
        "pf_last_error": (ctypes.c_char_p, []),
        "pf_device_sync": (i32, [vp]),
        "pf_eigs": (i32, [vp, ctypes.POINTER(PfEigStats)]),
        "pf_eigs_custom": (i32, [vp, vp, ctypes.POINTER(PfEigStats)]),
        "pf_eigs_copy": (i32, [vp, vp, i64]),
        "pf_generate": (i32, [vp, vp, u64, u64, i32, i32, dp, vp, i32]),
        "pf_matvec": (i32, [vp, i32, vp, vp, i32, i32]),
        "pf_cg": (i32, [vp, vp, vp, i32, ctypes.POINTER(PfCgOpts), ctypes.POINTER(PfCgStats)]),
        "pf_value_transform": (i32, [vp, i32, dp, vp, i64, i32]),
        "pf_philox_normals": (i32, [vp, u64, u64, u64, i64, vp, i32]),
        "pf_philox_raw": (i32, [vp, vp, vp, i64, vp, i32]),
    }
    sig.update({
        "pf_spectral_targets": (i32, [vp, i32, dp, vp]),
        "pf_mlp_create": (i32, [vp, i32, u64, ctypes.POINTER(vp)]),
        "pf_mlp_destroy": (i32, [vp]),
        "pf_mlp_nparams": (i32, [vp, ctypes.POINTER(i64)]),
        "pf_mlp_params": (i32, [vp, vp, i64, i32]),
        "pf_mlp_forward": (i32, [vp, vp, i32, vp]),
        "pf_mlp_gradients": (i32, [vp, vp, vp, i32, vp, ctypes.POINTER(dp)]),
        "pf_mlp_train": (i32, [vp, vp, vp, i32, i32, ctypes.c_float, ctypes.POINTER(dp)]),
        "pf_set_profiling": (i32, [vp, i32]),
        "pf_batch_create": (i32, [ctypes.POINTER(PfDesc), i32, ctypes.POINTER(vp)]),
        "pf_batch_destroy": (i32, [vp]),
        "pf_batch_offsets": (i32, [vp, vp]),
        "pf_batch_generate": (i32, [vp, u64, u64, i32, dp, vp, i32, i32, vp]),
        "pf_batch_set_branches": (i32, [vp, i32]),
        "pf_batch_set_field_timing": (i32, [vp, i32]),
        "pf_batch_field_times": (i32, [vp, ctypes.POINTER(dp)]),
        "pf_batch_autotune": (i32, [vp, u64, u64, i32, dp, vp, i32, ctypes.POINTER(i32),
                                    ctypes.POINTER(dp)]),
        "pf_last_timings": (i32, [vp, ctypes.c_char_p, i32, ctypes.POINTER(dp), i32,
                                  ctypes.POINTER(i32)]),
        "pf_autotune": (i32, [vp, i32, u64, i32, dp, vp, ctypes.c_char_p,
                              ctypes.POINTER(PfTuneResult)]),
        "pf_set_tuning": (i32, [vp, ctypes.POINTER(PfTuning)]),
        "pf_get_tuning": (i32, [vp, ctypes.POINTER(PfTuning)]),
    })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names every caller may bind (must all be exported by the library)."""
    return ["pf_plan_create", "pf_plan_destroy", "pf_plan_sizes", "pf_last_error",
            "pf_device_sync", "pf_eigs", "pf_eigs_custom", "pf_eigs_copy", "pf_generate",
            "pf_matvec", "pf_cg", "pf_value_transform", "pf_philox_normals", "pf_philox_raw",
            "pf_spectral_targets", "pf_mlp_create", "pf_mlp_destroy", "pf_mlp_nparams",
            "pf_mlp_params", "pf_mlp_forward", "pf_mlp_gradients", "pf_mlp_train",
            "pf_set_profiling", "pf_last_timings", "pf_batch_create", "pf_batch_destroy",
            "pf_batch_offsets", "pf_batch_generate", "pf_batch_set_branches",
            "pf_batch_autotune", "pf_batch_set_field_timing", "pf_batch_field_times"]


def check(status: int):
    if status == PF_OK:
        return
    msg = load().pf_last_error().decode(errors="replace")
    if status == PF_ENEG:
        raise NegativeEigenvalueError(msg)
    raise RuntimeError(f"parafields_amd error {status}: {msg}")


def _ptr(a) -> int:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()  # torch tensor (device memory)


def _mem(a) -> int:
    return MEM_HOST if isinstance(a, np.ndarray) else MEM_DEVICE


class Plan:
    """One covariance model on one grid: the state a reference Matrix holds."""

    def __init__(self, desc: PfDesc):
        self.lib = load()
        self.desc = desc
        h = ctypes.c_void_p()
        check(self.lib.pf_plan_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        s = PfSizes()
        check(self.lib.pf_plan_sizes(self.h, ctypes.byref(s)))
        self.sizes = s
        self.domain = int(s.domain_size)
        self.dtype = np.float64 if desc.precision == 0 else np.float32

    def close(self):
        if self.h:
            self.lib.pf_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eigs(self) -> dict:
        st = PfEigStats()
        check(self.lib.pf_eigs(self.h, ctypes.byref(st)))
        return st.as_dict()

    def eigs_stats_nothrow(self):
        st = PfEigStats()
        status = self.lib.pf_eigs(self.h, ctypes.byref(st))
        return status, st.as_dict()

    def eigs_custom(self, values: np.ndarray) -> dict:
        st = PfEigStats()
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(self.lib.pf_eigs_custom(self.h, v.ctypes.data, ctypes.byref(st)))
        return st.as_dict()

    def eig_values(self) -> np.ndarray:
        out = np.empty(int(self.sizes.eig_entries), dtype=np.float64)
        check(self.lib.pf_eigs_copy(self.h, out.ctypes.data, out.size))
        return out

    def generate(self, batch: int = 1, noise: Optional[np.ndarray] = None, seed: int = 0,
                 first: int = 0, transform: str = "none", param: float = 0.0, out=None):
        if out is None:
            out = np.empty((batch, self.domain), dtype=self.dtype)
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            # DFT field: one draw of noise_per_field normals gives two fields
            draws = (batch + 1) // 2 if self.desc.field_backend == FIELD_BACKENDS["dft"] else batch
            assert nz.size == draws * int(self.sizes.noise_per_field)
        check(self.lib.pf_generate(self.h, None if nz is None else nz.ctypes.data, seed, first,
                                   batch, TRANSFORMS[transform], param, _ptr(out), _mem(out)))
        return out

    def matvec(self, x, op: str = "multiply", batch: int = 1, out=None):
        if out is None:
            out = np.empty_like(x)
        check(self.lib.pf_matvec(self.h, OPS[op], _ptr(x), _ptr(out), batch, _mem(x)))
        return out

    def cg(self, b, mode: str = "reference", max_iterations: int = 0, rel_tol: float = 1e-8,
           bin_log_multipliers: Optional[np.ndarray] = None, x=None):
        if x is None:
            x = np.empty_like(b)
        o = PfCgOpts()
        o.mode = CG_MODES[mode]
        o.max_iterations = max_iterations
        o.rel_tol = rel_tol
        keep = None
        if bin_log_multipliers is not None:
            keep = np.ascontiguousarray(bin_log_multipliers, dtype=np.float32)
            o.bin_log_multipliers = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            o.nbins = keep.size
        st = PfCgStats()
        check(self.lib.pf_cg(self.h, _ptr(b), _ptr(x), _mem(b), ctypes.byref(o),
                             ctypes.byref(st)))
        return x, {"iterations": st.iterations, "rel_residual": st.rel_residual,
                   "seconds": st.seconds}

    def autotune(self, batch: int, out, seed: int = 0, transform: str = "none",
                 param: float = 0.0, cache: Optional[str] = None) -> dict:
        """Time the candidate launch choices of pf_generate(batch) into the
        device tensor `out` and keep the fastest (pf_autotune)."""
        r = PfTuneResult()
        check(self.lib.pf_autotune(self.h, batch, seed, TRANSFORMS[transform], param, _ptr(out),
                                   None if cache is None else cache.encode(), ctypes.byref(r)))
        return {"best": r.best.as_dict(), "ms": r.ms, "default_ms": r.default_ms,
                "candidates": r.candidates, "from_cache": bool(r.from_cache)}

    def tuning(self) -> dict:
        t = PfTuning()
        check(self.lib.pf_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = PfTuning()
        check(self.lib.pf_get_tuning(self.h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(self.lib.pf_set_tuning(self.h, ctypes.byref(t)))

    def value_transform(self, data, transform: str, param: float = 0.0):
        check(self.lib.pf_value_transform(self.h, TRANSFORMS[transform], param, _ptr(data),
                                          data.size if isinstance(data, np.ndarray)
                                          else data.numel(), _mem(data)))
        return data

    def philox_normals(self, seed: int, realization: int, first: int, n: int, out=None):
        if out is None:
            out = np.empty(n, dtype=np.float64)
        check(self.lib.pf_philox_normals(self.h, seed, realization, first, n, _ptr(out),
                                         _mem(out)))
        return out

    def philox_raw(self, ctr, key, nblocks: int):
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.empty(4 * nblocks, dtype=np.uint32)
        check(self.lib.pf_philox_raw(self.h, c.ctypes.data, k.ctypes.data, nblocks,
                                     out.ctypes.data, MEM_HOST))
        return out

    def sync(self):
        check(self.lib.pf_device_sync(self.h))

    def spectral_targets(self, nbins: int, floor_value: float) -> np.ndarray:
        out = np.empty(nbins, dtype=np.float64)
        check(self.lib.pf_spectral_targets(self.h, nbins, floor_value, out.ctypes.data))
        return out

    def set_profiling(self, on: bool):
        check(self.lib.pf_set_profiling(self.h, int(on)))

    def last_timings(self):
        names = ctypes.create_string_buffer(8192)
        ms = (ctypes.c_double * 256)()
        n = ctypes.c_int32()
        check(self.lib.pf_last_timings(self.h, names, 8192, ms, 256, ctypes.byref(n)))
        labels = names.value.decode().split(";")[: n.value]
        return list(zip(labels, [ms[i] for i in range(min(n.value, 256))]))


class Mlp:
    """The learned preconditioner's network: FP32, exact-GELU hidden layers,
    trained on the device by the library's own kernels."""

    def __init__(self, dims, seed: int = 0):
        self.lib = load()
        self.dims = list(dims)
        d = np.ascontiguousarray(self.dims, dtype=np.int32)
        h = ctypes.c_void_p()
        check(self.lib.pf_mlp_create(d.ctypes.data, len(self.dims), seed, ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int64()
        check(self.lib.pf_mlp_nparams(self.h, ctypes.byref(n)))
        self.nparams = n.value

    def close(self):
        if self.h:
            self.lib.pf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        check(self.lib.pf_mlp_params(self.h, out.ctypes.data, self.nparams, 0))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self.lib.pf_mlp_params(self.h, p.ctypes.data, self.nparams, 1))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], self.dims[-1]), dtype=np.float32)
        check(self.lib.pf_mlp_forward(self.h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        check(self.lib.pf_mlp_gradients(self.h, x.ctypes.data, y.ctypes.data, x.shape[0],
                                        g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def train(self, x, y, steps: int, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_double()
        check(self.lib.pf_mlp_train(self.h, x.ctypes.data, y.ctypes.data, x.shape[0], steps, lr,
                                    ctypes.byref(loss)))
        return loss.value


class Batch:
    """Uneven batch: many fields with their own descriptors, one launch."""

    def __init__(self, descs):
        self.lib = load()
        n = len(descs)
        arr = (PfDesc * n)(*descs)
        h = ctypes.c_void_p()
        check(self.lib.pf_batch_create(arr, n, ctypes.byref(h)))
        self.h = h
        self.n = n
        off = np.empty(n + 1, dtype=np.int64)
        check(self.lib.pf_batch_offsets(self.h, off.ctypes.data))
        self.offsets = off
        self.dtype = np.float32 if descs[0].precision == 1 else np.float64

    def close(self):
        if self.h:
            self.lib.pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate(self, seed=0, first=0, transform="none", param=0.0, out=None, use_graph=True,
                 allow_negative=False):
        if out is None:
            out = np.empty(int(self.offsets[-1]), dtype=self.dtype)
        st = (PfEigStats * self.n)()
        status = self.lib.pf_batch_generate(self.h, seed, first, TRANSFORMS[transform], param,
                                            _ptr(out), _mem(out), int(use_graph), st)
        if not (allow_negative and status == PF_ENEG):
            check(status)
        return out, [s.as_dict() for s in st]

    def set_branches(self, k: int):
        check(self.lib.pf_batch_set_branches(self.h, k))

    def autotune(self, seed=0, first=0, transform="none", param=0.0, out=None):
        if out is None:
            out = np.empty(int(self.offsets[-1]), dtype=self.dtype)
        best = ctypes.c_int32()
        ms = ctypes.c_double()
        check(self.lib.pf_batch_autotune(self.h, seed, first, TRANSFORMS[transform], param,
                                         _ptr(out), _mem(out), ctypes.byref(best),
                                         ctypes.byref(ms)))
        return best.value, ms.value

    def field_times(self, seed=0, first=0, transform="none", param=0.0, out=None):
        """Per-field device times (ms) of one eager serial pass (each field's
        eigenvalues and realization between two HIP events)."""
        if out is None:
            out = np.empty(int(self.offsets[-1]), dtype=self.dtype)
        check(self.lib.pf_batch_set_field_timing(self.h, 1))
        try:
            self.generate(seed=seed, first=first, transform=transform, param=param, out=out,
                          use_graph=False)
            ms = np.empty(len(self.offsets) - 1, dtype=np.float64)
            check(self.lib.pf_batch_field_times(self.h, ms.ctypes.data_as(
                ctypes.POINTER(ctypes.c_double))))
        finally:
            check(self.lib.pf_batch_set_field_timing(self.h, 0))
        return ms

    def field(self, out, i):
        return out[int(self.offsets[i]):int(self.offsets[i + 1])]
