"""Parity of the HIP path (through the C ABI of libparafields_amd.so) with the
FP64 CPU oracle on identical inputs. FP64 paths: <= 1e-11 relative; FP32
paths: <= 2e-5 of the field standard deviation; integer/bit work: exact."""
import ctypes

import numpy as np
import pytest

import np_helpers as nh
import oracle_py as orc
import pf_binding as pfb

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
FP32_TOL = 2e-5


def desc(dim, cells, **kw):
    kw.setdefault("extensions", [1.0] * dim)
    return pfb.make_desc(dim, cells, **kw)


def rel_max(got, want):
    """max |got - want| / max |want|; NaNs must sit at the same positions
    (the reference yields NaN where sqrt meets an unclamped negative)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    if not ok.any():
        return 0.0
    return float(np.abs(got[ok] - want[ok]).max() / max(np.abs(want[ok]).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(8)


def test_philox_raw_bit_exact():
    p = pfb.Plan(desc(2, [8, 8]))
    for ctr, key in [((0, 0, 0, 0), (0, 0)), ((5, 1, 7, 9), (0xDEADBEEF, 0x12345678)),
                     ((0xFFFFFFF0, 0xFFFFFFFF, 3, 0), (1, 2))]:
        got = p.philox_raw(ctr, key, 64)
        want = orc.philox_raw(ctr, key, 64)
        assert np.array_equal(got, want)


def test_philox_normals_fp64():
    p = pfb.Plan(desc(2, [8, 8]))
    got = p.philox_normals(7, 3, 11, 100000)
    want = orc.noise_philox(7, 3, 11, 100000)
    assert np.abs(got - want).max() < 1e-13


EIG_CASES = [
    (2, [256, 256], dict(covariance="exponential", corr_length=0.05)),  # cfg 1
    (2, [6, 5], dict(covariance="gaussian", corr_length=0.3, approximate=True)),
    (2, [24, 40], dict(covariance="gaussian", anisotropy="geometric",
                       corr_length=[0.1, 0.03, -0.02, 0.05], approximate=True)),
    (3, [16, 12, 8], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.1, 0.1, 0.05], variance=2.0, approximate=True)),
    (3, [8, 8, 12], dict(covariance="spherical", anisotropy="geometric",
                         corr_length=[0.2, 0.02, 0, 0.01, 0.1, 0.02, 0, 0.01, 0.08],
                         approximate=True)),
    (3, [10, 6, 8], dict(covariance="matern", matern_nu=0.8, corr_length=0.2,
                         approximate=True)),
    (3, [8, 8, 8], dict(covariance="gaussian", corr_length=0.5, approximate=True)),
]


@pytest.mark.parametrize("dim,cells,kw", EIG_CASES)
def test_eigs_match_oracle(dim, cells, kw):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    got = p.eig_values()
    assert status == wstatus
    assert rel_max(got, want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)
    assert st["smallest"] == pytest.approx(wst["smallest"], rel=1e-9, abs=1e-15)


# merge embedding (matrix.hh:652-775): DCT and R2C matrices, tensor / radial /
# radialPlus combinations (radial tapers need sqrt(d) < factor / 2), both
# sigmoids, default and explicit recursion counts
MERGE_CASES = [
    (2, [64, 48], dict(covariance="matern52", corr_length=0.25, factor=3,
                       sigmoid_combine="tensor")),
    (3, [16, 12, 8], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.3, 0.2, 0.25], factor=3)),
    (2, [32, 24], dict(covariance="exponential", corr_length=0.3, factor=4, sigmoid="smooth")),
    (2, [24, 20], dict(covariance="gaussian", anisotropy="geometric",
                       corr_length=[0.2, 0.1, 0.05, 0.3], factor=3, merge_recursions=3,
                       sigmoid_combine="radial")),
    (3, [8, 8, 6], dict(covariance="matern", matern_nu=1.2, corr_length=0.3, factor=3,
                        matrix="r2c")),
]


@pytest.mark.parametrize("dim,cells,kw", MERGE_CASES)
def test_merge_eigs_match_oracle(dim, cells, kw):
    d = desc(dim, cells, periodization="merge", approximate=True, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    assert status == wstatus
    assert rel_max(p.eig_values(), want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)


FOLD_CASES = [
    (2, [64, 48], dict(covariance="exponential", corr_length=0.2, factor=4)),
    (2, [32, 24], dict(covariance="matern52", corr_length=0.25, factor=3, sigmoid="smooth")),
    (3, [16, 12, 8], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.3, 0.2, 0.25], factor=4, fold_max_factor=1.5)),
    (2, [24, 20], dict(covariance="gaussian", anisotropy="geometric",
                       corr_length=[0.2, 0.1, 0.05, 0.3], factor=4, merge_recursions=3)),
    (3, [8, 8, 6], dict(covariance="matern", matern_nu=1.2, corr_length=0.3, factor=4,
                        matrix="r2c")),
]


@pytest.mark.parametrize("dim,cells,kw", FOLD_CASES)
def test_fold_eigs_match_oracle(dim, cells, kw):
    """Fold embedding: the device integrates the taper with composite
    Gauss-Legendre, the oracle in closed form (smoothstep) or by adaptive
    Simpson (smooth)."""
    d = desc(dim, cells, periodization="fold", approximate=True, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    assert status == wstatus
    assert rel_max(p.eig_values(), want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)


def test_fold_generate_matches_oracle():
    kw = dict(covariance="exponential", corr_length=0.2, factor=4, periodization="fold",
              approximate=True)
    d = desc(2, [64, 48], **kw)
    p = pfb.Plan(d)
    p.eigs()
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(31, npf)
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < FP64_TOL


COFOLD_CASES = [
    (2, [48, 40], dict(covariance="exponential", corr_length=0.3, factor=3)),
    (2, [32, 24], dict(covariance="matern52", corr_length=0.2, factor=3, sigmoid="smooth",
                       merge_recursions=-1)),
    (3, [12, 10, 8], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.3, 0.2, 0.25], factor=3, fold_max_factor=1.5,
                          merge_recursions=3)),
    (2, [24, 20], dict(covariance="gaussian", anisotropy="geometric",
                       corr_length=[0.2, 0.1, 0.05, 0.3], factor=3)),
    (3, [8, 8, 6], dict(covariance="exponential", corr_length=0.3, factor=4, matrix="r2c")),
]


@pytest.mark.parametrize("dim,cells,kw", COFOLD_CASES)
def test_cofold_eigs_match_oracle(dim, cells, kw):
    """Cofold: block minima / maxima on the device and their reduction on the
    host, then the remap with the device's Gauss-Legendre integral; the oracle
    reduces serially and integrates in closed form or by adaptive Simpson."""
    d = desc(dim, cells, periodization="cofold", approximate=True, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    assert status == wstatus
    assert rel_max(p.eig_values(), want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)


def test_cofold_generate_matches_oracle_and_batches_reject_it():
    d = desc(2, [48, 40], covariance="exponential", corr_length=0.3, factor=3,
             periodization="cofold", approximate=True)
    p = pfb.Plan(d)
    p.eigs()
    noise = orc.noise_cpp(37, int(p.sizes.noise_per_field))
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < FP64_TOL
    d32 = desc(2, [48, 40], covariance="exponential", corr_length=0.3, factor=3,
               periodization="cofold", approximate=True, precision="fp32")
    with pytest.raises(RuntimeError, match="cofold"):
        pfb.Batch([desc(2, [16, 16], precision="fp32"), d32])


def test_merge_generate_matches_oracle():
    """A field on the merge embedding (factor 3, no negative eigenvalues left)
    from host noise and from device Philox noise, FP64 and FP32."""
    kw = dict(covariance="matern52", corr_length=0.25, factor=3, periodization="merge",
              sigmoid_combine="tensor")
    d = desc(2, [64, 48], **kw)
    p = pfb.Plan(d)
    st = p.eigs()
    assert st["negative"] == 0
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(23, npf)
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < FP64_TOL
    p32 = pfb.Plan(desc(2, [64, 48], precision="fp32", **kw))
    got = p32.generate(1, seed=4)[0]
    want = orc.generate(d, orc.noise_philox(4, 0, 0, npf))
    assert np.abs(got - want).max() / want.std() < FP32_TOL


def test_merge_rejects_unbuilt_periodizations():
    with pytest.raises(RuntimeError):
        pfb.Plan(desc(2, [8, 8], covariance="gaussian", factor=3, periodization="merge"))
    d = desc(2, [8, 8], factor=3)
    d.periodization = 4
    with pytest.raises(RuntimeError):
        pfb.Plan(d)


R2C_CHECK_ALL_CASES = [
    (2, [48, 96], dict(covariance="gaussian", anisotropy="geometric",
                       corr_length=[0.08, 0.02, -0.01, 0.03], approximate=True)),
    (2, [1024, 64], dict(covariance="exponential", corr_length=0.05)),
]


@pytest.mark.parametrize("path", ["fast", "generic"])
@pytest.mark.parametrize("dim,cells,kw", R2C_CHECK_ALL_CASES)
def test_eigs_r2c_check_all_match_oracle(dim, cells, kw, path, monkeypatch):
    """R2C matrix with the full real-part check (the uneven batch's setting):
    the specialised 2D path (FP64 r2c along x, forward DFT along y, one fused
    classify/square-root pass) and the generic path both match the oracle."""
    if path == "generic":
        monkeypatch.setenv("PF_NO_FAST", "1")
    d = desc(dim, cells, matrix="r2c", r2c_check_all=True, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    assert status == wstatus
    assert rel_max(p.eig_values(), want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)
    # fields: FP64 bar, except where eigenvalues sit at the rounding level of
    # the transform (the Gaussian case has thousands at |lambda| ~ 1e-17 of
    # lambda_max); their square roots carry FFT rounding differences of
    # sqrt(1e-17) ~ 3e-9 into the field
    tol = 1e-7 if kw["covariance"] == "gaussian" else FP64_TOL
    noise = orc.noise_cpp(3, int(p.sizes.noise_per_field))
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < tol


@pytest.mark.parametrize("cells", [[24, 16], [48, 96]])
def test_eigs_custom_values_even_part_matches_full_transform(cells, monkeypatch):
    """Custom covariance values (pf_eigs_custom, the padded 2H x N1 array) on
    a 2D R2C plan with the full real-part check, for an arbitrary array that
    is not point-symmetric: the even-part path (values symmetrised on the
    host, half the rows, real y pass) gives the eigenvalues of numpy's 2D FFT
    (real part / N_ext, negatives clamped), and the eigenvalues and
    classification of the full-grid generic kernels (PF_NO_FAST=1)."""
    d = desc(2, cells, covariance="exponential", corr_length=0.1, matrix="r2c",
             r2c_check_all=True, approximate=True)
    N0, N1 = 2 * cells[0], 2 * cells[1]
    H = N0 // 2 + 1
    c = np.random.default_rng(5).standard_normal((N1, N0))
    c[0, 0] += 40.0
    vals = np.zeros((N1, 2 * H))
    vals[:, :N0] = c
    p = pfb.Plan(d)
    st = p.eigs_custom(vals.ravel())
    got = p.eig_values().reshape(N1, H)
    want = np.maximum(np.fft.fft2(c).real[:, :H] / (N0 * N1), 0.0)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    monkeypatch.setenv("PF_NO_FAST", "1")
    pg = pfb.Plan(d)
    stg = pg.eigs_custom(vals.ravel())
    assert np.abs(got - pg.eig_values().reshape(N1, H)).max() <= 1e-13 * np.abs(want).max()
    for k in ("small", "zero", "negative"):
        assert st[k] == stg[k], (k, st, stg)


def test_negative_eigenvalue_error():
    d = desc(2, [16, 16], covariance="gaussian", corr_length=0.5)
    p = pfb.Plan(d)
    with pytest.raises(pfb.NegativeEigenvalueError):
        p.generate(1, seed=1)


GEN_CASES = [
    (2, [256, 256], dict(covariance="exponential", corr_length=0.05)),  # cfg 1
    (2, [48, 20], dict(covariance="exponential", anisotropy="geometric",
                       corr_length=[0.1, 0.03, -0.02, 0.05], approximate=True)),
    (2, [10, 6], dict(covariance="cubic", corr_length=0.4, approximate=True)),
    (3, [16, 12, 8], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.1, 0.1, 0.05], variance=2.0, approximate=True)),
    (3, [8, 10, 6], dict(covariance="exponential", corr_length=0.2, transposed=0)),
    (3, [12, 8, 8], dict(covariance="matern52", anisotropy="geometric",
                         corr_length=[0.2, 0.02, 0, 0.01, 0.1, 0.02, 0, 0.01, 0.08],
                         approximate=True)),
]


def test_r2c_unchecked_half_gives_nan_like_reference():
    """R2C matrix + Gaussian: rounding-level negative eigenvalues in the second
    half of the complex buffer escape the interleaved check/clamp
    (r2cmatrixbackend.hh:378-400 read by matrix.hh:270-291), so sqrt yields
    NaN in generateField; the device path reproduces the oracle's NaNs."""
    d = desc(2, [48, 20], covariance="gaussian", anisotropy="geometric",
             corr_length=[0.1, 0.03, -0.02, 0.05], approximate=True)
    p = pfb.Plan(d)
    noise = orc.noise_cpp(42, int(p.sizes.noise_per_field))
    got = p.generate(1, noise=noise)[0]
    want = orc.generate(d, noise)
    assert np.isnan(want).all() and np.isnan(got).all()


@pytest.mark.parametrize("dim,cells,kw", GEN_CASES)
def test_generate_fp64_host_noise(dim, cells, kw):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    noise = orc.noise_cpp(42, 2 * int(p.sizes.noise_per_field))
    got = p.generate(2, noise=noise)
    npf = int(p.sizes.noise_per_field)
    for b in range(2):
        want = orc.generate(d, noise[b * npf:(b + 1) * npf])
        assert rel_max(got[b], want) < FP64_TOL


@pytest.mark.parametrize("dim,cells,kw", GEN_CASES[1:])
def test_generate_fp64_philox(dim, cells, kw):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    got = p.generate(2, seed=123456789012, first=5)
    for b in range(2):
        nz = orc.noise_philox(123456789012, 5 + b, 0, int(p.sizes.noise_per_field))
        want = orc.generate(d, nz)
        assert rel_max(got[b], want) < FP64_TOL


@pytest.mark.parametrize("dim,cells,kw", GEN_CASES[1:])
def test_generate_fp32_philox(dim, cells, kw):
    d = desc(dim, cells, precision="fp32", **kw)
    d64 = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    got = p.generate(3, seed=7)
    for b in range(3):
        nz = orc.noise_philox(7, b, 0, int(p.sizes.noise_per_field))
        want = orc.generate(d64, nz)
        assert np.abs(got[b] - want).max() / want.std() < FP32_TOL


@pytest.mark.parametrize("transform,param", [("logNormal", 0.0), ("foldedNormal", 0.0),
                                             ("sign", 0.0), ("sign", 0.5), ("boxCox", 0.3)])
def test_generate_value_transforms(transform, param):
    d = desc(3, [8, 8, 8], covariance="exponential", corr_length=0.2)
    p = pfb.Plan(d)
    noise = orc.noise_cpp(3, int(p.sizes.noise_per_field))
    raw = orc.generate(d, noise)
    want = orc.value_transform(raw, transform, param)
    got = p.generate(1, noise=noise, transform=transform, param=param)[0]
    if transform == "sign":
        # identical except where |raw - param| is at rounding level
        mask = np.abs(raw - param) > 1e-12
        assert np.array_equal(got[mask], want[mask])
    else:
        assert rel_max(got, want) < FP64_TOL


def test_generate_linear_in_noise():
    d = desc(3, [16, 16, 16], covariance="exponential", corr_length=0.1)
    p = pfb.Plan(d)
    n = int(p.sizes.noise_per_field)
    a = orc.noise_cpp(1, n)
    b = orc.noise_cpp(2, n)
    fa, fb, fab = (p.generate(1, noise=v)[0] for v in (a, b, 2 * a - 0.5 * b))
    assert rel_max(fab, 2 * fa - 0.5 * fb) < 1e-12


MV_CASES = [
    (2, [64, 48], dict(covariance="exponential", corr_length=0.1)),
    (2, [30, 20], dict(covariance="matern32", anisotropy="geometric",
                       corr_length=[0.1, 0.03, -0.02, 0.05], approximate=True)),
    (3, [16, 12, 10], dict(covariance="matern32", anisotropy="axiparallel",
                           corr_length=[0.1, 0.1, 0.05])),
    (3, [8, 8, 8], dict(covariance="exponential", corr_length=0.2, transposed=0)),
]


@pytest.mark.parametrize("dim,cells,kw", MV_CASES)
@pytest.mark.parametrize("op", ["multiply", "root", "inverse_embed"])
def test_matvec_fp64(dim, cells, kw, op):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    x = np.random.default_rng(5).standard_normal((2, p.domain))
    got = p.matvec(x, op=op, batch=2)
    for b in range(2):
        want = orc.matvec(d, x[b], op)
        assert rel_max(got[b], want) < (FP64_TOL if op != "inverse_embed" else 1e-9)


@pytest.mark.parametrize("op", ["multiply", "root"])
def test_matvec_fp32(op):
    d = desc(3, [16, 16, 16], covariance="exponential", corr_length=0.1, precision="fp32")
    d64 = desc(3, [16, 16, 16], covariance="exponential", corr_length=0.1)
    p = pfb.Plan(d)
    x = np.random.default_rng(6).standard_normal(p.domain)
    got = p.matvec(x.astype(np.float32), op=op)
    want = orc.matvec(d64, x, op)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-5


def test_multiply_inverse_reference_cg():
    cells = [12, 10, 8]
    d = desc(3, cells, covariance="exponential", corr_length=0.15)
    p = pfb.Plan(d)
    b = np.random.default_rng(7).standard_normal(p.domain)
    x, st = p.cg(b, mode="reference")
    want, it = orc.multiply_inverse(d, b)
    # innerCG stops when the energy change ratio falls below 1e-16
    # (matrix.hh:1421-1424), i.e. at rounding level: the iteration at which
    # that happens depends on the dot-product summation order, so the counts
    # of the two runs differ while both solutions sit at the converged answer.
    assert 0 < st["iterations"] <= 100 and 0 < it <= 100
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-7
    assert st["rel_residual"] < 1e-7


@pytest.mark.parametrize("mode", ["plain", "embed_pcg"])
def test_pcg_modes(mode):
    cells = [10, 8, 8]
    d = desc(3, cells, covariance="exponential", corr_length=0.2, nugget=1e-3,
             cg_iterations=2000)
    p = pfb.Plan(d)
    b = np.random.default_rng(8).standard_normal(p.domain)
    x, st = p.cg(b, mode=mode, rel_tol=1e-8)
    want, it, rr = orc.pcg(d, b, mode=mode, maxit=2000, rel_tol=1e-8)
    assert st["rel_residual"] <= 1e-8
    assert abs(st["iterations"] - it) <= 2
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-6


def test_value_transform_entry_point():
    p = pfb.Plan(desc(2, [8, 8]))
    v = np.linspace(-2, 2, 101)
    got = p.value_transform(v.copy(), "logNormal")
    assert rel_max(got, np.exp(v)) < 1e-15


def test_reference_smoke_matrix():
    # test/fieldtest.cc:26-57 in both precisions: generate() must not throw.
    for prec in ("fp64", "fp32"):
        for dim, cells in [(2, [16, 16]), (3, [8, 8, 8])]:
            for kind in ["exponential", "gaussian", "spherical"]:
                p = pfb.Plan(desc(dim, cells, covariance=kind, corr_length=0.05, precision=prec))
                f = p.generate(1, seed=11)
                assert np.all(np.isfinite(f))


def test_config2_scale_fp32_vs_fp64_oracle():
    """512^3 embedding (256^3 cells), Matern nu=1.5, axiparallel, one
    realization of Philox seed 7 in FP32 against the FP64 oracle."""
    kw = dict(covariance="matern32", variance=2.0, anisotropy="axiparallel",
              corr_length=[0.1, 0.1, 0.05], approximate=True)
    d32 = desc(3, [256, 256, 256], precision="fp32", **kw)
    d64 = desc(3, [256, 256, 256], **kw)
    p = pfb.Plan(d32)
    st = p.eigs()
    got = p.generate(1, seed=7, first=0)[0]
    orc.set_threads(16)
    nz = orc.noise_philox(7, 0, 0, int(p.sizes.noise_per_field))
    want = orc.generate(d64, nz)
    del nz
    err = np.abs(got - want).max() / want.std()
    l2 = nh.rel_l2(got, want)
    print(f"config-2 FP32 vs FP64 oracle: max-abs/std {err:.3e}, rel-L2 {l2:.3e}, "
          f"eig stats {st}", flush=True)
    assert err < FP32_TOL


CONFIG2_KW = dict(covariance="matern32", variance=2.0, anisotropy="axiparallel",
                  corr_length=[0.1, 0.1, 0.05], approximate=True)


def test_batch64_pipeline_matches_oracle_at_chunk_boundaries():
    """A batch of 64 runs the two-stream pipeline (8 chunks of 8 realizations).
    On a config-2-shaped field (3D 64^3 cells, 128^3 embedding, Matern-3/2,
    axiparallel, DCT matrix + R2C field, FP32), the first and last
    realizations of chunks (0, 7, 8, 63) match the FP64 oracle on identical
    Philox noise and equal the single-realization path bit for bit."""
    cells = [64, 64, 64]
    p = pfb.Plan(desc(3, cells, precision="fp32", **CONFIG2_KW))
    d64 = desc(3, cells, **CONFIG2_KW)
    npf = int(p.sizes.noise_per_field)
    out = p.generate(64, seed=7, first=1000)
    for r in (0, 7, 8, 63):
        single = p.generate(1, seed=7, first=1000 + r)[0]
        assert np.array_equal(out[r], single), r
        want = orc.generate(d64, orc.noise_philox(7, 1000 + r, 0, npf))
        err = np.abs(out[r] - want).max() / want.std()
        print(f"batch-64 realization {r}: max-abs/std {err:.3e}", flush=True)
        assert err < FP32_TOL, r


def test_config2_batch64_logNormal_full_size():
    """The bench workload itself: 64 FP32 realizations of the 256^3-cell field
    on the 512^3 embedding, log-normal transform, into device memory through
    the two-stream pipeline. Realizations 0, 8 and 63 equal the
    single-realization path bit for bit; log of realization 63 matches the FP64
    oracle's Gaussian field on identical Philox noise (seed 7)."""
    import torch
    cells = [256, 256, 256]
    p = pfb.Plan(desc(3, cells, precision="fp32", **CONFIG2_KW))
    p.eigs()
    out = torch.empty((64, p.domain), dtype=torch.float32, device="cuda")
    p.generate(64, seed=7, first=0, transform="logNormal", out=out)
    torch.cuda.synchronize()
    for r in (0, 8, 63):
        single = p.generate(1, seed=7, first=r, transform="logNormal")[0]
        assert np.array_equal(out[r].cpu().numpy(), single), r
    got = np.log(out[63].cpu().numpy().astype(np.float64))
    del out
    p.close()
    torch.cuda.empty_cache()
    orc.set_threads(16)
    d64 = desc(3, cells, **CONFIG2_KW)
    want = orc.generate(d64, orc.noise_philox(7, 63, 0, int(orc.sizes(d64).noise_per_field)))
    err = np.abs(got - want).max() / want.std()
    print(f"config-2 batch-64 realization 63 (log of logNormal output) vs FP64 oracle: "
          f"max-abs/std {err:.3e}, rel-L2 {nh.rel_l2(got, want):.3e}", flush=True)
    assert err < FP32_TOL


FASTPATH_CASES = [
    (2, [256, 256], dict(covariance="exponential", corr_length=0.05)),
    (3, [32, 48, 64], dict(covariance="matern32", anisotropy="axiparallel",
                           corr_length=[0.1, 0.1, 0.05], variance=2.0, approximate=True)),
    (3, [64, 32, 24], dict(covariance="exponential", corr_length=0.1, transposed=0)),
    (2, [96, 192], dict(covariance="matern52", anisotropy="geometric",
                        corr_length=[0.1, 0.03, -0.02, 0.05], approximate=True)),
    # 1024-point x pass (padded tiles of 4 FP32 / 2 FP64 lines) and 512-point
    (2, [1024, 48], dict(covariance="exponential", corr_length=0.05)),
    (2, [512, 96], dict(covariance="exponential", corr_length=0.08, transposed=0)),
]


@pytest.mark.parametrize("dim,cells,kw", FASTPATH_CASES)
@pytest.mark.parametrize("path", ["fast", "fast_half_tiles", "generic"])
def test_specialised_and_generic_kernels(dim, cells, kw, path, monkeypatch):
    """The compile-time specialised kernels (also with half-width tiles in the
    y and x passes, PF_TILE_DIV=2) and the generic kernels all match the oracle
    (PF_NO_FAST=1 routes every pass through the generic ones)."""
    if path == "generic":
        monkeypatch.setenv("PF_NO_FAST", "1")
    if path == "fast_half_tiles":
        monkeypatch.setenv("PF_TILE_DIV", "2")
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    noise = orc.noise_cpp(5, int(p.sizes.noise_per_field))
    want = orc.generate(d, noise)
    assert rel_max(p.generate(1, noise=noise)[0], want) < FP64_TOL
    got = p.generate(2, seed=99, first=3)
    for b in range(2):
        w = orc.generate(d, orc.noise_philox(99, 3 + b, 0, int(p.sizes.noise_per_field)))
        assert rel_max(got[b], w) < FP64_TOL
    d32 = desc(dim, cells, precision="fp32", **kw)
    p32 = pfb.Plan(d32)
    g32 = p32.generate(2, seed=99, first=3)
    for b in range(2):
        w = orc.generate(d, orc.noise_philox(99, 3 + b, 0, int(p.sizes.noise_per_field)))
        assert np.abs(g32[b] - w).max() / w.std() < FP32_TOL
    x = np.random.default_rng(1).standard_normal(p.domain)
    for op in ("multiply", "root", "inverse_embed"):
        assert rel_max(p.matvec(x, op=op), orc.matvec(d, x, op)) < (
            FP64_TOL if op != "inverse_embed" else 1e-9)


DCT_CASES = [
    (2, [8, 6], dict(covariance="exponential", corr_length=0.2)),               # 2D transposed
    (2, [16, 12], dict(covariance="gaussian", corr_length=0.1, transposed=0)),
    (3, [6, 4, 8], dict(covariance="exponential", corr_length=0.2)),            # 3D transposed
    (3, [8, 8, 6], dict(covariance="matern32", anisotropy="axiparallel",
                        corr_length=[0.2, 0.1, 0.15], transposed=0)),
    (3, [32, 16, 64], dict(covariance="matern52", anisotropy="geometric",
                           corr_length=[0.05, 0.01, 0, 0.01, 0.04, 0.01, 0, 0.01, 0.03],
                           approximate=True)),
    # every pass length in the specialised set (M = 16 ... 1024)
    (2, [64, 32], dict(covariance="matern32", corr_length=0.1, transposed=0)),
    (3, [32, 32, 16], dict(covariance="exponential", corr_length=0.1)),
]

# Pass length M = 1024 (FP32 tiles of 8 x 1024 complex, exactly 64 KB of LDS;
# FP64 4 x 1024) in the last-axis pass, the x pass and the middle pass; every
# pass of these shapes has a specialised kernel.
DCT_FAST_CASES = [
    (2, [16, 1024], dict(covariance="exponential", corr_length=0.05, transposed=0,
                         approximate=True)),
    (2, [1024, 16], dict(covariance="exponential", corr_length=0.05, approximate=True)),
    (3, [32, 1024, 16], dict(covariance="matern32", anisotropy="axiparallel",
                             corr_length=[0.1, 0.02, 0.1], transposed=0, approximate=True)),
]


@pytest.mark.parametrize("path", ["fast", "generic"])
@pytest.mark.parametrize("dim,cells,kw", DCT_CASES + DCT_FAST_CASES)
def test_dctdst_generate_fp64(dim, cells, kw, path, monkeypatch):
    """Host noise and Philox noise against the oracle, through the
    compile-time specialised DCT/DST passes (the default where the pass
    lengths have one) and through the generic ones (PF_NO_FAST=1)."""
    if path == "generic":
        monkeypatch.setenv("PF_NO_FAST", "1")
    d = desc(dim, cells, field="dctdst", matrix="dct", **kw)
    p = pfb.Plan(d)
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(17, 2 * npf)
    got = p.generate(2, noise=noise)
    for b in range(2):
        assert rel_max(got[b], orc.generate(d, noise[b * npf:(b + 1) * npf])) < FP64_TOL
    gp = p.generate(1, seed=5, first=2)[0]
    assert rel_max(gp, orc.generate(d, orc.noise_philox(5, 2, 0, npf))) < FP64_TOL


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("dim,cells,kw", DCT_FAST_CASES)
def test_dctdst_fast_passes_run(dim, cells, kw, precision):
    """The shapes of DCT_FAST_CASES go through the specialised passes (their
    profiling labels end in _fast) and the generic path is not taken."""
    p = pfb.Plan(desc(dim, cells, field="dctdst", matrix="dct", precision=precision, **kw))
    p.generate(1, seed=3)
    p.set_profiling(True)
    p.generate(2, seed=3)
    labels = [name for name, _ in p.last_timings()]
    p.set_profiling(False)
    want = ["dct_gen_last_fast"] + (["dct_mid_fast"] if dim == 3 else []) + ["dct_x_fast"]
    assert [x for x in labels if x.startswith("dct_")] == want, labels


@pytest.mark.parametrize("dim,cells,kw", DCT_CASES[2:] + DCT_FAST_CASES)
def test_dctdst_generate_fp32(dim, cells, kw):
    d = desc(dim, cells, field="dctdst", matrix="dct", **kw)
    d32 = desc(dim, cells, field="dctdst", matrix="dct", precision="fp32", **kw)
    p = pfb.Plan(d32)
    npf = int(p.sizes.noise_per_field)
    got = p.generate(2, seed=8, transform="sign", param=0.5)
    raw = p.generate(2, seed=8)
    for b in range(2):
        want = orc.generate(d, orc.noise_philox(8, b, 0, npf))
        assert np.abs(raw[b] - want).max() / want.std() < FP32_TOL
        mask = np.abs(want - 0.5) > 1e-4 * want.std()
        assert np.array_equal(got[b][mask], np.where(want < 0.5, -1.0, 1.0)[mask])


def test_dctdst_rejects_factor_other_than_2():
    with pytest.raises(RuntimeError):
        pfb.Plan(desc(3, [4, 4, 4], field="dctdst", matrix="dct", factor=3))


def test_cpp_host_api_end_to_end(tmp_path):
    """The C++ drop-in API (include/parafields_amd/randomfield.hh) driven by
    pf_host_driver: reference smoke matrix, then generate(42) /
    timesMatrix / timesMatrixRoot / generate(43) / timesInverseMatrix on the
    2D 256x256 exponential field, each against the oracle."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pfb.LIB_PATH), "pf_host_driver")
    r = subprocess.run([exe, "smoke"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "smoke: ok" in r.stdout, r.stdout + r.stderr
    pre = str(tmp_path / "f")
    r = subprocess.run([exe, "field", pre], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = desc(2, [256, 256], covariance="exponential", corr_length=0.05)
    npf = int(orc.sizes(d).noise_per_field)
    gen = orc.generate(d, orc.noise_cpp(42, npf))
    load = lambda k: np.fromfile(f"{pre}.{k}.bin", dtype=np.float64)
    assert rel_max(load("gen"), gen) < FP64_TOL
    mul = orc.matvec(d, gen)
    assert rel_max(load("mul"), mul) < FP64_TOL
    assert rel_max(load("root"), orc.matvec(d, mul, "root")) < FP64_TOL
    g43 = orc.generate(d, orc.noise_cpp(43, npf))
    assert rel_max(load("gen43"), g43) < FP64_TOL
    inv, _ = orc.multiply_inverse(d, g43)
    assert np.linalg.norm(load("inv") - inv) / np.linalg.norm(inv) < 1e-7


def test_cpp_trend_components_end_to_end(tmp_path):
    """RandomField with mean, slope, disk and block trend components
    (pf_host_driver trend): the stochastic part is the oracle's field of the
    first samples of the libstdc++ seed-42 stream; the 12 trend coefficients
    are the next 12 samples times the square roots of their variances, in
    component order (randomfield.hh:550-554, trend.hh:324-342); bulkEvaluate
    adds the trend with coefficients mean + shift at the cell centres (the
    component shapes themselves are pinned by tests/host/trend_check.cc)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pfb.LIB_PATH), "pf_host_driver")
    pre = str(tmp_path / "t")
    r = subprocess.run([exe, "trend", pre], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "trend: ok 12" in r.stdout, r.stdout + r.stderr
    d = desc(2, [64, 48], covariance="exponential", corr_length=0.1)
    npf = int(orc.sizes(d).noise_per_field)
    stream = orc.noise_cpp(42, npf + 12)
    load = lambda k: np.fromfile(f"{pre}.{k}.bin", dtype=np.float64)
    stoch = load("stoch")
    assert rel_max(stoch, orc.generate(d, stream[:npf])) < FP64_TOL
    # mean | slope x, y | disk centre x, y, radius, value | block centre x, y, extent x, y, value
    mean = np.array([0.5, 1, -2, 0.3, 0.6, 0.1, 1, 0.5, 0.5, 0.2, 0.3, 2])
    var = np.array([0.25, 0.5, 2, 0.01, 0.01, 0.001, 0.5, 0.01, 0.01, 0, 0, 1])
    shift = load("shift")
    np.testing.assert_allclose(shift, stream[npf:] * np.sqrt(var), rtol=1e-15, atol=0)
    c = mean + shift
    x = (np.arange(64) + 0.5) / 64
    y = ((np.arange(48) + 0.5) / 48)[:, None]
    trend = (c[0] + c[1] * (x - 0.5) + c[2] * (y - 0.5)
             + np.exp(-((x - c[3]) ** 2 + (y - c[4]) ** 2) / c[5] ** 2 * c[6])
             + np.where((np.abs(x - c[7]) <= 0.5 * c[9]) & (np.abs(y - c[8]) <= 0.5 * c[10]),
                        c[11], 0.0))
    np.testing.assert_allclose(load("bulk").reshape(48, 64) - stoch.reshape(48, 64), trend,
                               rtol=1e-12, atol=1e-12)


# ------------------------------------------------ learned preconditioner --

def test_mlp_gradients_match_fp64_oracle():
    dims = [6, 64, 64, 64, 64]
    m = pfb.Mlp(dims, seed=3)
    p = m.params()
    rng = np.random.default_rng(4)
    x = rng.normal(size=(40, 6)).astype(np.float32)
    y = rng.normal(size=(40, 64)).astype(np.float32)
    g, loss = m.gradients(x, y)
    wl, wg, wout = orc.mlp(dims, p.astype(np.float64), x, y)
    assert abs(loss - wl) / wl < 1e-5
    np.testing.assert_allclose(m.forward(x), wout, rtol=0, atol=1e-4 * np.abs(wout).max())
    assert np.linalg.norm(g - wg) / np.linalg.norm(wg) < 1e-4
    assert np.abs(g - wg).max() < 1e-3 * np.abs(wg).max()


def test_mlp_training_fits_a_smooth_map():
    dims = [6, 64, 64, 64, 8]
    m = pfb.Mlp(dims, seed=1)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(256, 6)).astype(np.float32)
    y = np.stack([np.sin(x[:, 0] * (k + 1)) + x[:, 1] * x[:, 2] for k in range(8)], 1)
    l0 = m.gradients(x, y.astype(np.float32))[1]
    l1 = m.train(x, y.astype(np.float32), 1500, 3e-3)
    assert l1 < 0.02 * l0


def test_spectral_targets_match_oracle():
    d = desc(3, [16, 12, 8], covariance="exponential", corr_length=0.2, nugget=1e-3)
    p = pfb.Plan(d)
    floor = 1e-3 / p.sizes.extended_size
    assert rel_max(p.spectral_targets(32, floor), orc.spectral_targets(d, 32, floor)) < 1e-12


def test_learned_pcg_matches_oracle():
    cells = [16, 12, 10]
    d = desc(3, cells, covariance="exponential", corr_length=0.2, nugget=1e-3,
             cg_iterations=5000)
    p = pfb.Plan(d)
    t = orc.spectral_targets(d, 24, 1e-3 / p.sizes.extended_size)
    b = np.random.default_rng(12).standard_normal(p.domain)
    x, st = p.cg(b, mode="learned_pcg", rel_tol=1e-8, bin_log_multipliers=t.astype(np.float32))
    want, it, rr = orc.pcg_learned(d, b, t.astype(np.float32).astype(np.float64), 5000, 1e-8)
    assert st["rel_residual"] <= 1e-8 and rr <= 1e-8
    assert abs(st["iterations"] - it) <= 2
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-6


def test_trained_network_gains_over_plain_cg():
    """Train the MLP on method-generated targets (small grid), then solve
    held-out cases: learned PCG reaches 1e-8 in far fewer iterations than
    plain CG (the fallback), both judged by the FP64 true residual."""
    import learned_precond as lp
    cases = lp.sample_cases(160, seed=5, cells=24)
    x, y = lp.build_training_set(cases)
    model, losses = lp.train(x, y, steps=1800, lr=2e-3)
    assert losses[-1] < 0.05
    held = lp.sample_cases(4, seed=99, cells=24)
    rng = np.random.default_rng(11)
    gains = []
    for c in held:
        b = rng.standard_normal(24 ** 3)
        _, sp = lp.solve(c, b, "plain")
        _, sl = lp.solve(c, b, "learned_pcg", model)
        assert sl["rel_residual"] <= 1e-8 and sp["rel_residual"] <= 1e-8
        gains.append(sp["iterations"] / sl["iterations"])
    print("learned-PCG iteration gains over plain CG:", gains, flush=True)
    assert min(gains) >= 3.0


def _uneven_descs(n, seed, precision="fp32"):
    """Config-3 style fields: 2D, sides from {64..1024 as 2^k and 3*2^k} (small
    here), Gaussian covariance with full geometric anisotropy, R2C matrix with
    the full real-part check (the reference's interleaved check turns
    Gaussian R2C fields into NaN)."""
    rng = np.random.default_rng(seed)
    sides = [16, 24, 32, 48, 64, 96, 128]
    out = []
    for _ in range(n):
        n0, n1 = (int(rng.choice(sides)) for _ in range(2))
        th, r, L = rng.uniform(0, np.pi), rng.uniform(1, 8), rng.uniform(0.02, 0.2)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        A = R @ np.diag([L, L / r])
        out.append(desc(2, [n0, n1], covariance="gaussian", anisotropy="geometric",
                        corr_length=list(A.ravel()), approximate=True, r2c_check_all=True,
                        precision=precision))
    return out


def test_uneven_batch_graph_matches_single_plans():
    descs = _uneven_descs(14, seed=3)
    descs.append(desc(3, [16, 16, 24], covariance="matern32", anisotropy="axiparallel",
                      corr_length=[0.1, 0.1, 0.05], approximate=True, precision="fp32"))
    descs.append(desc(3, [8, 8, 8], covariance="exponential", corr_length=0.2,
                      field="dctdst", matrix="dct", precision="fp32"))
    b = pfb.Batch(descs)
    out_g, st = b.generate(seed=21, first=100, use_graph=True)
    out_g2, _ = b.generate(seed=21, first=100, use_graph=True)
    out_e, _ = b.generate(seed=21, first=100, use_graph=False)
    assert np.array_equal(out_g, out_g2) and np.array_equal(out_g, out_e)
    for i, d in enumerate(descs):
        p = pfb.Plan(d)
        single = p.generate(1, seed=21, first=100 + i)[0]
        assert np.array_equal(b.field(out_g, i), single), i
        assert np.all(np.isfinite(single))
    for i in (0, 5, 14, 15):
        d64 = descs[i]
        d64.precision = 0
        want = orc.generate(d64, orc.noise_philox(21, 100 + i, 0,
                                                  int(orc.sizes(d64).noise_per_field)))
        assert np.abs(b.field(out_g, i) - want).max() / want.std() < FP32_TOL


def test_uneven_batch_mixes_periodizations():
    """Fields of one covariance family with classical, merge and fold
    periodization in one batch: the covariance launches are grouped per family
    and periodization, so each field keeps the embedding of its own plan."""
    descs = []
    for i, (n0, n1) in enumerate([(32, 24), (48, 32), (24, 24), (64, 48), (16, 32), (40, 24)]):
        per = ("classical", "merge", "fold")[i % 3]
        descs.append(desc(2, [n0, n1], covariance="matern52", corr_length=0.2, factor=3,
                          periodization=per, approximate=True, precision="fp32"))
    b = pfb.Batch(descs)
    out, _ = b.generate(seed=5, first=40, use_graph=True)
    for i, d in enumerate(descs):
        single = pfb.Plan(d).generate(1, seed=5, first=40 + i)[0]
        assert np.array_equal(b.field(out, i), single), i
    for i in (0, 1, 2):
        d64 = descs[i]
        d64.precision = 0
        want = orc.generate(d64, orc.noise_philox(5, 40 + i, 0,
                                                  int(orc.sizes(d64).noise_per_field)))
        assert np.abs(b.field(out, i) - want).max() / want.std() < FP32_TOL


def test_uneven_batch_branches_and_autotune():
    """Forked graph branches (fields bin-packed over concurrent streams) give
    the same bits as the serial graph; the autotuner picks one of the
    candidate branch counts and keeps the result unchanged."""
    descs = _uneven_descs(40, seed=5)
    b = pfb.Batch(descs)
    ref, _ = b.generate(seed=9, first=3, use_graph=True)
    for k in (2, 4, 16):
        b.set_branches(k)
        out, _ = b.generate(seed=9, first=3, use_graph=True)
        out2, _ = b.generate(seed=9, first=3, use_graph=True)
        assert np.array_equal(out, ref), k
        assert np.array_equal(out2, ref), k
    best, ms = b.autotune(seed=9, first=3)
    assert best in (1, 2, 4, 8, 16) and ms > 0
    out, _ = b.generate(seed=9, first=3, use_graph=True)
    assert np.array_equal(out, ref)


def test_uneven_batch_groups_match_single_fields(monkeypatch):
    """Fields of one shape run as one group (one launch per kernel over the
    group, per-field covariance, threshold, Philox realization, root table and
    output offset): bit-identical to the same batch with every field through
    its own plan (PF_BATCH_GROUP=0), with the per-field eigenvalue statistics
    unchanged; one field per group is checked against the oracle."""
    rng = np.random.default_rng(17)
    kws = []
    for i in range(24):
        n0, n1 = [(64, 64), (96, 128), (128, 96), (64, 192)][i % 4]
        th, r, L = rng.uniform(0, np.pi), rng.uniform(1, 8), rng.uniform(0.02, 0.2)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        A = R @ np.diag([L, L / r])
        kws.append(dict(cells=[n0, n1], covariance="gaussian", anisotropy="geometric",
                        corr_length=list(A.ravel()), approximate=True, r2c_check_all=True,
                        threshold=1e-14 * (1 + i % 3)))
    descs = [desc(2, precision="fp32", **kw) for kw in kws]
    got, st = pfb.Batch(descs).generate(seed=5, first=7, transform="logNormal", use_graph=True)
    monkeypatch.setenv("PF_BATCH_GROUP", "0")
    ref, st_ref = pfb.Batch(descs).generate(seed=5, first=7, transform="logNormal",
                                            use_graph=True)
    assert np.array_equal(got, ref)
    assert st == st_ref
    monkeypatch.delenv("PF_BATCH_GROUP")
    b = pfb.Batch(descs)
    gauss, _ = b.generate(seed=5, first=7)
    for i in range(4):
        d64 = desc(2, **kws[i])
        n = int(pfb.Plan(d64).sizes.noise_per_field)
        want = orc.generate(d64, orc.noise_philox(5, 7 + i, 0, n))
        assert np.abs(b.field(gauss, i) - want).max() / want.std() < FP32_TOL


def test_uneven_batch_field_times():
    """Per-field device times of a timed eager pass: one positive time per
    field, the timed pass writes the same bits as the graph, and asking for
    times after a graph replay is a state error."""
    descs = _uneven_descs(6, seed=12)
    b = pfb.Batch(descs)
    ms = b.field_times(seed=2, first=1)          # first call: allocating pass, then the timed one
    assert ms.shape == (6,) and np.all(ms > 0) and np.all(np.isfinite(ms))
    ref, _ = b.generate(seed=2, first=1, use_graph=True)
    out = np.empty_like(ref)
    ms2 = b.field_times(seed=2, first=1, out=out)
    assert np.all(ms2 > 0) and np.array_equal(out, ref)
    b.generate(seed=2, first=1, use_graph=True)
    buf = np.empty(6)
    with pytest.raises(RuntimeError):
        pfb.check(b.lib.pf_batch_field_times(b.h, buf.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))


def test_uneven_batch_more_branches_than_fields():
    """A batch of 3 fields asked for 16 branches runs on 3 (no empty branch)
    and keeps the bits of the serial graph; the autotuner stops at 3."""
    descs = _uneven_descs(3, seed=8)
    b = pfb.Batch(descs)
    ref, _ = b.generate(seed=4, first=0, use_graph=True)
    b.set_branches(16)
    out, _ = b.generate(seed=4, first=0, use_graph=True)
    assert np.array_equal(out, ref)
    best, _ = b.autotune(seed=4, first=0)
    assert best in (1, 2)
    out, _ = b.generate(seed=4, first=0, use_graph=True)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("precision,matrix", [("fp32", "dct"), ("fp64", "r2c")])
def test_philox_fields_mean_and_covariance(precision, matrix):
    """Throughput mode statistics: 40000 device-Philox realizations of a small
    2D field have the generator's mean (0) and covariance A·Aᵀ, where A is the
    oracle's noise→field map (its columns are the fields of unit noise
    vectors; it carries the reference's kx-plane variance deficit). Every
    entry must sit within 6 standard errors of the sample covariance of a
    Gaussian vector, Var(S_ij) = (C_ii C_jj + C_ij^2) / B."""
    d = desc(2, [10, 8], covariance="exponential", corr_length=0.3, variance=1.5,
             matrix=matrix, precision=precision)
    d64 = desc(2, [10, 8], covariance="exponential", corr_length=0.3, variance=1.5,
               matrix=matrix)
    npf = int(orc.sizes(d64).noise_per_field)
    A = np.stack([orc.generate(d64, np.eye(npf)[i]) for i in range(npf)], axis=1)
    C = A @ A.T
    p = pfb.Plan(d)
    B = 40000
    f = p.generate(B, seed=2024, first=0).astype(np.float64)
    mean = f.mean(axis=0)
    S = np.cov(f, rowvar=False, bias=True) + np.outer(mean, mean)  # E[x xᵀ] about the true mean 0
    sd = np.sqrt(np.diag(C))
    assert np.all(np.abs(mean) < 6 * sd / np.sqrt(B))
    se = np.sqrt((np.outer(np.diag(C), np.diag(C)) + C ** 2) / B)
    assert np.all(np.abs(S - C) < 6 * se)
    # the deficit is real: the model covariance differs from A·Aᵀ on the diagonal
    assert abs(np.diag(C).mean() - 1.5) > 6 * np.sqrt(2 * 1.5 ** 2 / (B * len(sd)))


def test_gpumatrix_adapter_end_to_end(tmp_path):
    """The reference-side adapter (include/parafields_amd/gpumatrix.hh) run
    through reference-shaped traits (tests/host/gpumatrix_check.cc): every
    Matrix method against the oracle on the same libstdc++ noise, including
    refinement, geometric anisotropy (R2C matrix) and the DCT/DST field with
    the reference's accumulation (a second generate adds onto the field;
    operator* and multiplyRoot return input + product)."""
    import os
    import subprocess
    import test_host_api as tha
    exe = tha._build_gpumatrix_check(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, "run", str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr

    def rd(name):
        return np.fromfile(str(out / f"{name}.bin"), dtype=np.float64)

    base = dict(variance=1.5, approximate=True)
    d = desc(2, [40, 24], extensions=[1.0, 0.6], covariance="exponential", corr_length=0.1,
             **base)
    n = int(orc.sizes(d).noise_per_field)
    g = orc.generate(d, orc.noise_cpp(42, n))
    assert rel_max(rd("iso_gen"), g) < FP64_TOL
    assert rel_max(rd("iso_mult"), orc.matvec(d, g)) < FP64_TOL
    assert rel_max(rd("iso_root"), orc.matvec(d, g, "root")) < FP64_TOL
    want_inv, _ = orc.multiply_inverse(d, g)
    assert np.linalg.norm(rd("iso_inv") - want_inv) / np.linalg.norm(want_inv) < 1e-7
    assert np.array_equal(rd("iso_uncorr"), orc.noise_cpp(7, 40 * 24))
    assert np.all(rd("iso_var") == 1.5)
    d2 = desc(2, [80, 48], extensions=[1.0, 0.6], covariance="exponential", corr_length=0.1,
              **base)
    assert rel_max(rd("iso_refined_gen"),
                   orc.generate(d2, orc.noise_cpp(43, int(orc.sizes(d2).noise_per_field)))) < FP64_TOL
    da = desc(3, [12, 10, 8], covariance="exponential", anisotropy="geometric",
              corr_length=[0.1, 0.02, 0, 0.01, 0.08, 0, 0, 0, 0.05], matrix="r2c", **base)
    ga = orc.generate(da, orc.noise_cpp(5, int(orc.sizes(da).noise_per_field)))
    assert rel_max(rd("aniso_gen"), ga) < FP64_TOL
    assert rel_max(rd("aniso_mult"), orc.matvec(da, ga)) < FP64_TOL
    d1 = desc(1, [500], covariance="exponential", corr_length=0.05, **base)
    n1 = int(orc.sizes(d1).noise_per_field)
    z1 = orc.noise_cpp(11, 2 * n1)
    w1, w2 = orc.generate(d1, z1[:n1]), orc.generate(d1, z1[n1:])
    assert rel_max(rd("dft1_gen1"), w1[0]) < FP64_TOL
    assert rel_max(rd("dft1_gen2"), w1[1]) < FP64_TOL  # the spare field, no draw
    assert rel_max(rd("dft1_gen3"), w2[0]) < FP64_TOL
    assert rel_max(rd("dft1_mult"), orc.matvec(d1, w2[0])) < FP64_TOL
    want_inv1, _ = orc.multiply_inverse(d1, w2[0])
    assert np.linalg.norm(rd("dft1_inv") - want_inv1) / np.linalg.norm(want_inv1) < 1e-7
    dd = desc(3, [16, 12, 10], covariance="exponential", anisotropy="axiparallel",
              corr_length=[0.2, 0.1, 0.15], matrix="dct", field="dctdst", **base)
    nd = int(orc.sizes(dd).noise_per_field)
    z = orc.noise_cpp(9, 2 * nd)
    g1 = orc.generate(dd, z[:nd])
    g2 = g1 + orc.generate(dd, z[nd:])
    assert rel_max(rd("dct_gen1"), g1) < FP64_TOL
    assert rel_max(rd("dct_gen2"), g2) < FP64_TOL
    assert rel_max(rd("dct_mult"), g2 + orc.matvec(dd, g2)) < FP64_TOL
    assert rel_max(rd("dct_root"), g2 + orc.matvec(dd, g2, "root")) < FP64_TOL
    want_dinv, _ = orc.multiply_inverse(dd, g2)
    assert np.linalg.norm(rd("dct_inv") - want_dinv) / np.linalg.norm(want_dinv) < 1e-7
    assert np.all(np.isfinite(rd("custom_f32_gen"))) and np.all(np.isfinite(rd("custom_f32_inv")))


@pytest.mark.parametrize("dim,cells,kw", [
    (2, [24, 16], dict(covariance="exponential", corr_length=0.1)),
    (3, [16, 12, 10], dict(covariance="exponential", anisotropy="axiparallel",
                           corr_length=[0.2, 0.1, 0.15])),
    (3, [8, 12, 16], dict(covariance="matern52", corr_length=0.1, transposed=0)),
])
@pytest.mark.parametrize("op", ["multiply", "root", "inverse_embed"])
def test_dctdst_matvec_matches_oracle(dim, cells, kw, op):
    """pf_matvec on a DCT/DST-field plan against the oracle's restatement of
    the reference's per-type DCT/DST route (FP64 <= 1e-11; FP32 <= 2e-5 of
    the product's max)."""
    d = desc(dim, cells, field="dctdst", matrix="dct", approximate=True, **kw)
    x = np.random.default_rng(3).standard_normal(int(np.prod(cells)))
    want = orc.matvec(d, x, op)
    assert rel_max(pfb.Plan(d).matvec(x, op=op), want) < (FP64_TOL if op != "inverse_embed"
                                                          else 1e-9)
    d32 = desc(dim, cells, field="dctdst", matrix="dct", approximate=True, precision="fp32", **kw)
    if op != "inverse_embed":
        got = pfb.Plan(d32).matvec(x.astype(np.float32), op=op)
        assert np.abs(got - want).max() / np.abs(want).max() < FP32_TOL


def test_dctdst_multiply_inverse_matches_oracle():
    d = desc(3, [12, 10, 8], covariance="exponential", corr_length=0.2, field="dctdst",
             matrix="dct", cg_iterations=5000)
    b = np.random.default_rng(8).standard_normal(12 * 10 * 8)
    want, _ = orc.multiply_inverse(d, b)
    got, st = pfb.Plan(d).cg(b, mode="reference")
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-7
    for mode in ("plain", "embed_pcg"):
        x, st = pfb.Plan(d).cg(b, mode=mode, rel_tol=1e-10, max_iterations=5000)
        assert st["rel_residual"] <= 1e-10, mode
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-6, mode


def test_cpp_field_files_caches_refinement(tmp_path):
    """The C++ RandomField beyond generation (tests/host/field_check.cc): the
    .stoch.h5 / .trend / .field files written and read back, the cached
    inverse products of timesMatrix / timesInverseMatrix / timesMatrixRoot /
    timesInvMatRoot (randomfield.hh:1181-1310), timesInvMatRoot without a
    cache, refinement with a valid cache (2^d C_fine refine(C^{-1} x)) and
    without (cell copies), and the arithmetic, against the oracle."""
    import os
    import subprocess
    import test_host_api as tha
    inc, libs = tha._hdf5_link_args(tmp_path)
    exe = str(tmp_path / "field_check")
    libdir = os.path.dirname(pfb.LIB_PATH)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), *inc,
                           os.path.join(root, "tests", "host", "field_check.cc"), "-o", exe, *libs,
                           "-L", libdir, "-lparafields_amd", "-Wl,-rpath," + libdir])
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "all OK" in r.stdout, r.stdout + r.stderr

    def rd(name):
        return np.fromfile(str(out / f"{name}.bin"), dtype=np.float64)

    d = desc(2, [32, 24], extensions=[1.0, 0.75], covariance="exponential", corr_length=0.15,
             variance=1.2)
    n = int(orc.sizes(d).noise_per_field)
    z42 = orc.noise_cpp(42, n + 3)
    x = orc.generate(d, z42[:n])
    assert rel_max(rd("gen42"), x) < FP64_TOL
    assert rel_max(rd("cx"), orc.matvec(d, x)) < FP64_TOL
    assert rel_max(rd("cache_root"), orc.matvec(d, x, "root")) < FP64_TOL
    assert rel_max(rd("root_x"), orc.matvec(d, x, "root")) < FP64_TOL
    inv, _ = orc.multiply_inverse(d, x)
    want = orc.matvec(d, inv, "root")
    assert np.linalg.norm(rd("invroot_x") - want) / np.linalg.norm(want) < 1e-7
    # cell copies of the device's own field (bit for bit)
    fine = np.kron(rd("gen42").reshape(24, 32), np.ones((2, 2))).ravel()
    assert np.array_equal(rd("refined_plain"), fine)
    assert np.array_equal(rd("refined_cache"), 4 * fine)
    df = desc(2, [64, 48], extensions=[1.0, 0.75], covariance="exponential", corr_length=0.15,
              variance=1.2)
    assert rel_max(rd("refined_cached"), 4 * orc.matvec(df, fine)) < FP64_TOL
    z7 = orc.noise_cpp(7, n + 3)
    y = orc.generate(d, z7[:n])
    assert rel_max(rd("axpy"), x + 2 * y) < FP64_TOL
    sd = np.sqrt([0.25, 0.5, 0.5])
    dot = float(open(out / "dot.txt").read())
    want_dot = x @ y + (z42[n:] * sd) @ (z7[n:] * sd)
    assert abs(dot - want_dot) <= 1e-10 * abs(want_dot)
    # 1-D (DFT matrix and field): the second generate returns the spare field
    d1 = desc(1, [300], covariance="exponential", corr_length=0.05)
    n1 = int(orc.sizes(d1).noise_per_field)
    w42 = orc.generate(d1, orc.noise_cpp(42, n1))
    w44 = orc.generate(d1, orc.noise_cpp(44, n1))
    assert rel_max(rd("d1_gen42"), w42[0]) < FP64_TOL
    assert rel_max(rd("d1_gen43"), w42[1]) < FP64_TOL
    assert rel_max(rd("d1_gen44"), w44[0]) < FP64_TOL
    assert rel_max(rd("d1_mult44"), orc.matvec(d1, w44[0])) < FP64_TOL
