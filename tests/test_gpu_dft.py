"""Parity of the DFT backends on the GPU with the FP64 oracle: the DFT matrix
(dftmatrixbackend.hh; every entry of the grid, classification over all of
them), the DFT field (dftfieldbackend.hh; lambda (xi_1 + i xi_2) through the
backward complex transform, the real part as the field and the imaginary part
as the spare field), 1-D fields (the reference's defaults, matrix.hh:1660-1736;
one tile or a four-step split) and the products and CG on them."""
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
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(8)


# (dim, cells, kw): 1-D one tile, 1-D four-step (6000 = 2000 x 3 at FP64,
# 3000 x 2 at FP32), 1-D 2^20 cells (2048 x 1024), 2D and 3D both orders
CASES = [
    (1, [12], dict(covariance="exponential", corr_length=0.2)),
    (1, [3000], dict(covariance="matern32", corr_length=0.05, variance=2.0)),
    (1, [1 << 20], dict(covariance="exponential", corr_length=0.01)),
    (1, [500], dict(covariance="exponential", anisotropy="geometric", corr_length=[0.3])),
    (2, [40, 24], dict(covariance="exponential", anisotropy="geometric",
                       corr_length=[0.1, 0.02, 0.0, 0.08], matrix="dft", field="dft")),
    (2, [32, 20], dict(covariance="exponential", corr_length=0.1, matrix="dft", field="dft",
                       transposed=0)),
    (3, [12, 8, 10], dict(covariance="matern32", anisotropy="axiparallel",
                          corr_length=[0.2, 0.1, 0.15], matrix="dft", field="dft")),
    (3, [8, 10, 6], dict(covariance="exponential", corr_length=0.2, matrix="dft", field="dft",
                         transposed=0)),
]


# spectra with eigenvalues at rounding level (the 1-D spherical covariance's
# exact zeros, the Gaussian's tail): sqrt(lambda) of a value of 1e-17 moves by
# 1e-9 under rounding, so the fields of such cases only agree to that level
# with any other FFT (FFTW's included); eigenvalues and classification only
EIG_ONLY = [(1, [500], dict(covariance="spherical", anisotropy="geometric", corr_length=[0.3],
                            approximate=True)),
            (2, [40, 24], dict(covariance="gaussian", anisotropy="geometric",
                               corr_length=[0.1, 0.02, 0.0, 0.08], approximate=True,
                               matrix="dft", field="dft"))]


@pytest.mark.parametrize("dim,cells,kw", CASES + EIG_ONLY)
def test_dft_eigs_match_oracle(dim, cells, kw):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    want, wst, wstatus = orc.eigs(d, allow_negative=True)
    assert int(p.sizes.eig_entries) == int(p.sizes.extended_size) == want.size
    got = p.eig_values()
    assert status == wstatus
    assert rel_max(got, want) < FP64_TOL
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)
    p.close()


@pytest.mark.parametrize("dim,cells,kw", CASES)
def test_dft_generate_fp64_host_noise(dim, cells, kw):
    """Three fields from two draws: field 0 / 1 are the real / imaginary part
    of draw 0, field 2 the real part of draw 1."""
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(42, 2 * npf)
    got = p.generate(3, noise=noise)
    for j in range(2):
        want = orc.generate(d, noise[j * npf:(j + 1) * npf])
        for part in range(2):
            f = 2 * j + part
            if f < 3:
                assert rel_max(got[f], want[part]) < FP64_TOL, (j, part)
    p.close()


@pytest.mark.parametrize("dim,cells,kw", [CASES[1], CASES[2], CASES[4], CASES[6]])
@pytest.mark.parametrize("first,batch", [(0, 2), (3, 4), (5, 1)])
def test_dft_generate_philox(dim, cells, kw, first, batch):
    """Field f of the Philox stream is part f % 2 of draw f / 2 (Philox
    realization f / 2), whatever the first field and batch of the call."""
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    npf = int(p.sizes.noise_per_field)
    got = p.generate(batch, seed=987654321, first=first)
    cache = {}
    for i in range(batch):
        f = first + i
        if f // 2 not in cache:
            cache[f // 2] = orc.generate(d, orc.noise_philox(987654321, f // 2, 0, npf))
        assert rel_max(got[i], cache[f // 2][f % 2]) < FP64_TOL, f
    p.close()


@pytest.mark.parametrize("dim,cells,kw", [CASES[1], CASES[3], CASES[4], CASES[6]])
def test_dft_generate_fp32(dim, cells, kw):
    d = desc(dim, cells, precision="fp32", **kw)
    d64 = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    got = p.generate(2, seed=11, first=0)
    want = orc.generate(d64, orc.noise_philox(11, 0, 0, int(p.sizes.noise_per_field)))
    for part in range(2):
        err = np.abs(got[part].astype(np.float64) - want[part]).max() / want[part].std()
        assert err < FP32_TOL, (part, err)
    p.close()


@pytest.mark.parametrize("dim,cells,kw", [CASES[0], CASES[1], CASES[2], CASES[5], CASES[7]])
@pytest.mark.parametrize("op", ["multiply", "root", "inverse_embed"])
def test_dft_matvec_fp64(dim, cells, kw, op):
    d = desc(dim, cells, **kw)
    p = pfb.Plan(d)
    n = int(np.prod(cells))
    x = np.random.default_rng(2).standard_normal((2, n))
    got = p.matvec(x, op, batch=2)
    tol = FP64_TOL
    if op == "inverse_embed":
        # 1 / lambda: the error of any two FFTs grows with the spectrum's
        # condition number kappa (both agree to ~kappa * eps)
        lam, _, _ = orc.eigs(d)
        tol = max(tol, np.finfo(float).eps * lam.max() / lam.min())
    for b in range(2):
        assert rel_max(got[b], orc.matvec(d, x[b], op)) < tol
    p.close()


def test_dft_matvec_fp32_1d_four_step():
    kw = CASES[1][2]
    d = desc(1, [3000], precision="fp32", **kw)
    d64 = desc(1, [3000], **kw)
    p = pfb.Plan(d)
    x = np.random.default_rng(3).standard_normal(3000)
    got = p.matvec(x.astype(np.float32), "multiply")
    want = orc.matvec(d64, x)
    assert rel_max(got, want) < 1e-5
    p.close()


def test_dft_1d_multiply_inverse_and_pcg():
    """innerCG (PF_CG_REFERENCE) and plain / embedding PCG on a 1-D field
    against the dense solve."""
    cells = [400]
    d = desc(1, cells, covariance="exponential", corr_length=0.05, nugget=1e-3,
             cg_iterations=500)
    p = pfb.Plan(d)
    s = p.sizes
    C = nh.dense_cov(cells, [1.0], "exponential", "none", [0.05])
    b = np.random.default_rng(6).standard_normal(400)
    want = np.linalg.solve((C + 1e-3 * np.eye(400)) / s.extended_size, b)
    for mode in ("plain", "embed_pcg"):
        x, st = p.cg(b, mode=mode, max_iterations=5000, rel_tol=1e-10)
        assert st["rel_residual"] <= 1e-10
        assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-7, mode
    d2 = desc(1, cells, covariance="exponential", corr_length=0.05, cg_iterations=500)
    p2 = pfb.Plan(d2)
    C2 = C / s.extended_size
    x, st = p2.cg(b, mode="reference")
    want2 = np.linalg.solve(C2, b)
    assert np.linalg.norm(x - want2) / np.linalg.norm(want2) < 1e-6
    p.close()
    p2.close()


@pytest.mark.parametrize("dim,cells,kw", [
    (2, [48, 32], dict(covariance="exponential", anisotropy="geometric",
                       corr_length=[0.1, 0.02, 0.0, 0.08])),
    (3, [12, 8, 10], dict(covariance="matern52", anisotropy="axiparallel",
                          corr_length=[0.2, 0.1, 0.15])),
])
def test_dft_matrix_with_r2c_field(dim, cells, kw):
    """Matrix<T, DFTMatrixBackend> in 2D/3D (the default field is R2C): the
    DFT matrix's classification over all entries, the R2C field's noise and
    products."""
    d = desc(dim, cells, matrix="dft", field="r2c", **kw)
    p = pfb.Plan(d)
    status, st = p.eigs_stats_nothrow()
    _, wst, _ = orc.eigs(d, allow_negative=True)
    for k in ("small", "zero", "negative", "entries"):
        assert st[k] == wst[k], (k, st, wst)
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(7, npf)
    assert rel_max(p.generate(1, noise=noise)[0], orc.generate(d, noise)) < FP64_TOL
    x = np.random.default_rng(1).standard_normal(int(np.prod(cells)))
    assert rel_max(p.matvec(x, "multiply"), orc.matvec(d, x, "multiply")) < FP64_TOL
    p.close()


def test_dct_matrix_with_dft_field():
    """Matrix<T, DCTMatrixBackend, DFTFieldBackend>: DCTMatrixBackend::eval
    folds the full-grid frequencies onto its half array."""
    d = desc(2, [36, 28], covariance="matern32", corr_length=0.1, matrix="dct", field="dft")
    p = pfb.Plan(d)
    npf = int(p.sizes.noise_per_field)
    noise = orc.noise_cpp(3, npf)
    got = p.generate(2, noise=noise)
    want = orc.generate(d, noise)
    assert rel_max(got[0], want[0]) < FP64_TOL and rel_max(got[1], want[1]) < FP64_TOL
    p.close()


def test_dft_statistics_of_the_spare_field():
    """Full-size property: 64 fields (32 draws) of a 1-D 2^16-cell field;
    both parts have the model variance and the real and imaginary parts of one
    draw are uncorrelated."""
    d = desc(1, [1 << 16], covariance="exponential", corr_length=0.02, precision="fp32")
    p = pfb.Plan(d)
    f = p.generate(64, seed=5, first=0).astype(np.float64)
    var = f.var()
    assert abs(var - 1.0) < 0.05, var
    re, im = f[0::2], f[1::2]
    corr = np.mean([np.corrcoef(re[j], im[j])[0, 1] for j in range(32)])
    assert abs(corr) < 0.05, corr
    p.close()


def test_dft_rejected_in_batches_and_1d_rules():
    with pytest.raises(RuntimeError):
        pfb.Batch([desc(1, [64])])
    with pytest.raises(RuntimeError):
        pfb.Plan(desc(1, [64], matrix="dft", field="r2c"))
    with pytest.raises(RuntimeError):
        pfb.Plan(desc(2, [16, 16], matrix="r2c", field="dft"))
