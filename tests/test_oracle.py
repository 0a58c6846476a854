"""Pins the FP64 CPU oracle (oracle/pf_oracle.cc) before it is trusted as the
checker of the HIP path. Each check is independent of the oracle's own code:
textbook closed forms, scipy/numpy library routines (fftn, dctn/dstn type 1,
eigvalsh, solve, kv), brute-force dense matrices, and statistical identities
of the generated fields."""
import numpy as np
import pytest
from scipy import fft as sfft
from scipy import special

import np_helpers as nh
import oracle_py as orc
import pf_binding as pfb


def desc(dim, cells, **kw):
    kw.setdefault("extensions", [1.0] * dim)
    return pfb.make_desc(dim, cells, **kw)


# ---------------------------------------------------------------- noise --

def test_cpp_noise_golden_seed42():
    # libstdc++ default_random_engine(42) + normal_distribution<double>; values
    # measured for the reference's CppRNGBackend (cpprngbackend.hh:57-100).
    v = orc.noise_cpp(42, 2)
    assert v[0] == -1.7141127395876619
    assert v[1] == 0.17805682714513033


def test_philox_known_answers():
    # Random123 philox4x32_10 known-answer vectors (Salmon et al. 2011).
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = orc.philox_raw(ctr, key, 1)
        assert tuple(int(x) for x in got) == want


def test_philox_normals_moments():
    x = orc.noise_philox(7, 3, 0, 400000)
    assert abs(x.mean()) < 5 / np.sqrt(x.size)
    assert abs(x.var() - 1) < 6 * np.sqrt(2 / x.size)
    # pairs from one Box-Muller draw are uncorrelated
    assert abs(np.mean(x[0::2] * x[1::2])) < 6 / np.sqrt(x.size / 2)
    # counter layout: a later window equals the tail of a longer draw
    y = orc.noise_philox(7, 3, 5, 11)
    assert np.array_equal(y, x[5:16])
    # realizations are distinct streams
    z = orc.noise_philox(7, 4, 0, 16)
    assert not np.allclose(z, x[:16])


# ------------------------------------------------------------ covariance --

@pytest.mark.parametrize("nu", [0.3, 0.5, 1.0, 1.5, 2.5, 3.7])
def test_bessel_k_vs_scipy(nu):
    for x in [1e-3, 0.1, 0.9, 1.99, 2.01, 5.0, 30.0]:
        want = special.kv(nu, x)
        assert abs(orc.bessel_k(nu, x) - want) <= 1e-12 * abs(want)


@pytest.mark.parametrize("kind", ["exponential", "gaussian", "spherical", "matern32",
                                  "matern52", "matern", "cauchy", "generalizedCauchy",
                                  "gammaExponential", "dampedOscillation", "cubic",
                                  "separableExponential"])
@pytest.mark.parametrize("aniso", ["none", "axiparallel", "geometric"])
def test_covariance_closed_forms(kind, aniso):
    dim = 2
    rng = np.random.default_rng(1)
    corr = {"none": [0.3], "axiparallel": [0.3, 0.1],
            "geometric": [0.3, 0.05, -0.02, 0.1]}[aniso]
    kw = dict(matern_nu=1.7, exp_gamma=1.3, cauchy_alpha=1.5, cauchy_beta=0.8)
    d = desc(dim, [8, 8], covariance=kind, variance=2.0, corr_length=corr, anisotropy=aniso,
             **kw)
    pts = rng.uniform(-0.5, 0.5, size=(200, dim))
    got = orc.cov_eval(d, pts)
    want = nh.cov_at(pts, kind, aniso, corr, var=2.0, nu=1.7, gamma=1.3, alpha=1.5, beta=0.8)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


def test_covariance_3d_geometric_and_spherical():
    rng = np.random.default_rng(2)
    A = [0.2, 0.03, 0.01, -0.02, 0.1, 0.0, 0.01, 0.02, 0.05]
    for kind in ["spherical", "matern52"]:
        d = desc(3, [4, 4, 4], covariance=kind, corr_length=A, anisotropy="geometric")
        pts = rng.uniform(-0.3, 0.3, size=(100, 3))
        np.testing.assert_allclose(orc.cov_eval(d, pts), nh.cov_at(pts, kind, "geometric", A),
                                   rtol=1e-12, atol=1e-14)


def test_cubic_rejected_in_3d():
    d = desc(3, [4, 4, 4], covariance="cubic", corr_length=0.2)
    with pytest.raises(RuntimeError):
        orc.eigs(d)


# ----------------------------------------------------------- eigenvalues --

def to_natural_dct(v, d):
    s = orc.sizes(d)
    P = [s.extended[i] // 2 + 1 for i in range(s.dim)]
    if s.dim == 2:
        arr = v.reshape(P[1], P[0])
        return arr.T.T if not s.transposed else v.reshape(P[0], P[1]).T
    if s.transposed:
        return v.reshape(P[1], P[2], P[0]).transpose(1, 0, 2)
    return v.reshape(P[2], P[1], P[0])


@pytest.mark.parametrize("dim,cells,kind,aniso,corr", [
    (2, [8, 6], "exponential", "none", [0.2]),
    (2, [16, 16], "gaussian", "axiparallel", [0.1, 0.05]),
    (3, [4, 6, 4], "matern32", "axiparallel", [0.1, 0.1, 0.05]),
    (3, [4, 4, 4], "spherical", "none", [0.3]),
])
def test_eigs_dct_vs_library(dim, cells, kind, aniso, corr):
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, approximate=True)
    v, st, _ = orc.eigs(d, allow_negative=True)
    nat = to_natural_dct(v, d)
    full = nh.full_eigs(cells, [1.0] * dim, kind, aniso, corr)
    assert np.abs(full.imag).max() < 1e-12
    P = [c + 1 for c in cells]
    half = full.real[tuple(slice(0, p) for p in P[::-1])]
    # scipy's DCT-I is FFTW's REDFT00
    g = nh.wrapped_grid(cells, [1.0] * dim)
    c = nh.cov_at(g.reshape(-1, dim), kind, aniso, corr).reshape(g.shape[:-1])
    dct = sfft.dctn(c[tuple(slice(0, p) for p in P[::-1])], type=1) / c.size
    expect = np.where(half < 0, 0.0, half)  # clamping of negatives (matrix.hh:289-290)
    np.testing.assert_allclose(nat, expect, rtol=0, atol=1e-13 * np.abs(half).max())
    np.testing.assert_allclose(np.where(dct < 0, 0, dct), expect, atol=1e-13 * np.abs(half).max())
    # classification counts (matrix.hh:270-291) from the unclamped values
    thr = 1e-14
    assert st["negative"] == int(np.sum(half <= -thr))
    assert st["zero"] == int(np.sum(np.abs(half) < thr))
    assert st["small"] == int(np.sum((half >= thr) & (half < 1e-6)))
    assert st["smallest"] == pytest.approx(half.min(), abs=1e-13 * np.abs(half).max())


def test_eigs_are_circulant_spectrum():
    # Eigenvalues of the dense circulant embedding matrix (eigvalsh) equal
    # N_ext times the oracle's eigenvalues, expanded by even symmetry.
    cells, corr = [4, 3], [0.4]
    d = desc(2, cells, covariance="exponential", corr_length=corr)
    v, _, _ = orc.eigs(d)
    P = [c + 1 for c in cells]
    N = [2 * c for c in cells]
    half = v.reshape(P[1], P[0])
    full = np.empty((N[1], N[0]))
    for ky in range(N[1]):
        for kx in range(N[0]):
            full[ky, kx] = half[ky if ky < P[1] else N[1] - ky, kx if kx < P[0] else N[0] - kx]
    g = nh.wrapped_grid(cells, [1.0, 1.0])
    pts = g.reshape(-1, 2)
    # circulant: entry (i, j) = c(wrap(x_i - x_j)) on the periodic extended grid
    Lx, Ly = 2.0, 2.0
    dlt = pts[:, None, :] - pts[None, :, :]
    dlt[..., 0] = np.where(dlt[..., 0] > 0.5 * Lx, dlt[..., 0] - Lx, dlt[..., 0])
    dlt[..., 0] = np.where(dlt[..., 0] < -0.5 * Lx, dlt[..., 0] + Lx, dlt[..., 0])
    dlt[..., 1] = np.where(dlt[..., 1] > 0.5 * Ly, dlt[..., 1] - Ly, dlt[..., 1])
    dlt[..., 1] = np.where(dlt[..., 1] < -0.5 * Ly, dlt[..., 1] + Ly, dlt[..., 1])
    C = nh.cov_at(dlt.reshape(-1, 2), "exponential", "none", corr).reshape(len(pts), len(pts))
    ev = np.linalg.eigvalsh(C)
    np.testing.assert_allclose(np.sort(full.ravel() * full.size), ev, atol=1e-12)


@pytest.mark.parametrize("dim,cells,corr", [
    (2, [8, 6], [0.2, 0.05, -0.03, 0.1]),
    (3, [4, 4, 6], [0.2, 0.02, 0.0, 0.01, 0.1, 0.02, 0.0, 0.01, 0.08]),
])
def test_eigs_r2c_vs_library(dim, cells, corr):
    d = desc(dim, cells, covariance="gaussian", corr_length=corr, anisotropy="geometric",
             approximate=True)
    assert d.matrix_backend == pfb.MATRIX_BACKENDS["r2c"]
    v, st, _ = orc.eigs(d, allow_negative=True)
    s = orc.sizes(d)
    N = [2 * c for c in cells]
    H = N[0] // 2 + 1
    spec = nh.full_eigs(cells, [1.0] * dim, "gaussian", "geometric", corr)[..., :H]
    # storage order of the reference (3D: transposed, kx fastest then kz then ky)
    store = spec if dim == 2 else spec.transpose(1, 0, 2)
    store = np.ascontiguousarray(store)
    # interleaved classification over the first H*N1*N2 reals (matrix.hh:270-291)
    reals = store.view(np.float64).ravel().copy()
    count = store.size
    head = reals[:count]
    thr = 1e-14
    assert st["negative"] == int(np.sum(head <= -thr))
    assert st["zero"] == int(np.sum(np.abs(head) < thr))
    assert st["small"] == int(np.sum((head >= thr) & (head < 1e-6)))
    head[head < 0] = 0.0
    reals[:count] = head
    want = reals.view(np.complex128).real
    np.testing.assert_allclose(v, want, atol=1e-13 * np.abs(want).max())
    assert s.transposed == (dim == 3)


def test_negative_eigenvalues_raise_unless_approximate():
    # Gaussian with a long correlation length and factor 2 embeds badly.
    d = desc(2, [16, 16], covariance="gaussian", corr_length=0.5)
    with pytest.raises(pfb.NegativeEigenvalueError):
        orc.generate(d, np.zeros(orc.sizes(d).noise_per_field))
    d.approximate = 1
    v, st, status = orc.eigs(d)
    assert status == 0 and st["negative"] > 0 and v.min() >= 0


def test_reference_smoke_matrix_does_not_throw():
    # test/fieldtest.cc:26-57: 2D 16x16 and 3D 8^3, corrLength 0.05, three families.
    for dim, cells in [(2, [16, 16]), (3, [8, 8, 8])]:
        for kind in ["exponential", "gaussian", "spherical"]:
            d = desc(dim, cells, covariance=kind, corr_length=0.05)
            n = orc.sizes(d).noise_per_field
            f = orc.generate(d, orc.noise_cpp(1, n))
            assert np.all(np.isfinite(f))


# ------------------------------------------------------------ generation --

def linear_map(d, nsamp):
    """Field as a linear map of the noise: columns = fields of unit noise."""
    cols = []
    for j in range(nsamp):
        z = np.zeros(nsamp)
        z[j] = 1.0
        cols.append(orc.generate(d, z))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,matrix", [
    (2, [4, 3], "exponential", [0.3], "none", "dct"),
    (2, [4, 4], "gaussian", [0.2, 0.05, 0.0, 0.15], "geometric", "r2c"),
    (3, [2, 3, 2], "exponential", [0.3], "none", "dct"),
    (3, [2, 2, 2], "matern52", [0.3, 0.2, 0.1], "axiparallel", "r2c"),
    (2, [4, 3], "gaussian", [0.2, 0.05, 0.0, 0.15], "geometric", "dft"),
    (3, [2, 3, 2], "matern32", [0.3, 0.2, 0.25], "axiparallel", "dft"),
])
def test_generate_r2c_covariance_identity(dim, cells, kind, corr, aniso, matrix):
    """L L^T = C - D, where D is the variance lost on the kx in {0, N0/2}
    planes because their +-k entries get independent noise (the reference's
    c2r keeps only the Hermitian part of those planes)."""
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, matrix=matrix,
             approximate=True)
    s = orc.sizes(d)
    L = linear_map(d, int(s.noise_per_field))
    C = nh.dense_cov(cells, [1.0] * dim, kind, aniso, corr)
    full = nh.full_eigs(cells, [1.0] * dim, kind, aniso, corr).real
    full = np.maximum(full, 0.0) if matrix in ("dct", "dft") else full
    N = [2 * c for c in cells]
    x = nh.cell_coords(cells, [1.0] * dim) / (np.array([1.0] * dim) / np.array(cells))  # int idx
    D = np.zeros_like(C)
    it = np.ndindex(*full.shape)
    for kk in it:
        k = kk[::-1]  # (kx, ky[, kz])
        if k[0] not in (0, N[0] // 2):
            continue
        if all((2 * k[i]) % N[i] == 0 for i in range(dim)):
            continue
        ph = 2 * np.pi * sum(k[i] * x[:, i] / N[i] for i in range(dim))
        D += 0.5 * full[kk] * np.cos(ph[:, None] - ph[None, :])
    if matrix == "r2c":
        # R2C eigen: full spectrum as stored (no clamping needed: approximate with
        # the interleaved clamp would only touch negative values)
        assert full.min() > -1e-12
    np.testing.assert_allclose(L @ L.T, C - D, atol=1e-12)


def test_generate_single_mode_noise_order_3d():
    """One unit sample at the xi_1 slot of storage entry s produces
    w * a * cos(2 pi k.x / N), with k the reference's transposed 3D order
    (kx fastest, then kz, then ky; matrix.hh:374-404, r2cfieldbackend.hh:190-199)."""
    cells = [2, 3, 4]
    d = desc(3, cells, covariance="exponential", corr_length=0.3)
    s = orc.sizes(d)
    N = [2 * c for c in cells]
    H = N[0] // 2 + 1
    full = nh.full_eigs(cells, [1.0] * 3, "exponential", "none", [0.3]).real
    x = np.stack(np.meshgrid(*[np.arange(c) for c in cells[::-1]], indexing="ij")[::-1],
                 axis=-1).reshape(-1, 3)
    for sidx in [0, 1, 2, 7, H + 1, H * N[2] + 3, H * N[2] * 2 + H * 3 + 2, 2 * H * N[2] - 1]:
        kx, kz, ky = sidx % H, (sidx // H) % N[2], sidx // (H * N[2])
        m = full[kz, ky, kx]
        selfconj = all((2 * k) % n == 0 for k, n in zip((kx, ky, kz), N))
        w = 1.0 if kx in (0, N[0] // 2) else 2.0
        ph = 2 * np.pi * (kx * x[:, 0] / N[0] + ky * x[:, 1] / N[1] + kz * x[:, 2] / N[2])
        for slot in (0, 1):
            z = np.zeros(s.noise_per_field)
            z[2 * sidx + slot] = 1.0
            f = orc.generate(d, z)
            if selfconj:
                want = np.sqrt(m) * np.cos(ph) if slot == 0 else 0 * ph
            else:
                a = np.sqrt(m) / np.sqrt(2)
                want = w * a * (np.cos(ph) if slot == 0 else -np.sin(ph))
            np.testing.assert_allclose(f, want, atol=1e-13)


def _dctdst_single_sample_cases():
    cases = []
    for dim, cells in ((2, [3, 5]), (3, [2, 3, 4])):
        for tr in (0, 1):
            for t in range(1 << dim):
                cases.append((dim, cells, tr, t))
    return cases


@pytest.mark.parametrize("dim,cells,tr,typ", _dctdst_single_sample_cases())
def test_generate_dctdst_single_sample_closed_form(dim, cells, tr, typ):
    """One unit sample at storage index s of parity type t gives the closed form
    2^{-d/2} 2^{border/2} sqrt(m_k) prod_i B_i(k_i, x_i), with B = c cos(2 pi k x / N)
    on even axes (c = 1 on the boundary k in {0, N/2}, else 2: inverse REDFT00) and
    B = 2 sin(2 pi k x / N) on odd axes (inverse RODFT00 on the interior; a boundary
    index of an odd axis gives the zero field). Type bit i is the parity of axis i
    (x = bit 0; dctdstfieldbackend.hh:149-165); the samples of type t are the
    t-th block of prod P_i in the noise stream (matrix.hh:342-372); within a block
    the storage order is x fastest with the two last axes swapped when the
    transform is transposed (3D: x, z, y; 2D: y, x; dctdstfieldbackend.hh:859-872).
    Distinct P_i per axis make a swapped order or a reversed bit order fail."""
    d = desc(dim, cells, covariance="exponential", corr_length=[0.15], field="dctdst",
             matrix="dct", transposed=tr, approximate=True)
    s = orc.sizes(d)
    N = [2 * c for c in cells]
    P = [n // 2 + 1 for n in N]
    count = int(np.prod(P))
    assert s.noise_per_field == count << dim
    full = nh.full_eigs(cells, [1.0] * dim, "exponential", "none", [0.15]).real
    x = nh.cell_coords(cells, [float(c) for c in cells]).round().astype(int)  # integer cell index
    order = list(range(dim))  # storage axes, fastest first
    if tr:
        order[-1], order[-2] = order[-2], order[-1]
    odd = [(typ >> i) & 1 for i in range(dim)]
    rng = np.random.default_rng(100 + typ + 10 * tr)
    picks = {0, count - 1, 1, P[order[0]], P[order[0]] + 1}
    picks |= set(int(v) for v in rng.integers(0, count, 6))
    for sidx in sorted(picks):
        k = [0] * dim
        r = sidx
        for ax in order:
            k[ax] = r % P[ax]
            r //= P[ax]
        m = full[tuple(k[::-1])]
        assert m > 0
        want = np.full(x.shape[0], np.sqrt(m) * 2.0 ** (-dim / 2))
        zero = False
        border = 0
        for i in range(dim):
            edge = k[i] in (0, P[i] - 1)
            ph = 2 * np.pi * k[i] * x[:, i] / N[i]
            if odd[i]:
                zero |= edge
                want = want * 2 * np.sin(ph)
            else:
                border += edge
                want = want * (1.0 if edge else 2.0) * np.cos(ph)
        want = 0 * want if zero else want * 2.0 ** (border / 2)
        z = np.zeros(s.noise_per_field)
        z[typ * count + sidx] = 1.0
        f = orc.generate(d, z)
        np.testing.assert_allclose(f, want, atol=1e-13, err_msg=f"s={sidx} k={k}")


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,tr", [
    (2, [3, 4], "exponential", [0.3], "none", -1),
    (2, [4, 4], "gaussian", [0.2, 0.1], "axiparallel", 0),
    (3, [2, 3, 2], "exponential", [0.3], "none", -1),
    (3, [3, 2, 2], "matern32", [0.2, 0.1, 0.3], "axiparallel", 0),
])
def test_generate_dctdst_covariance_identity(dim, cells, kind, corr, aniso, tr):
    """The DCT/DST backend reproduces the embedded covariance exactly:
    L L^T = C on the original grid."""
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, field="dctdst",
             matrix="dct", transposed=tr, approximate=True)
    s = orc.sizes(d)
    L = linear_map(d, int(s.noise_per_field))
    C = nh.dense_cov(cells, [1.0] * dim, kind, aniso, corr)
    np.testing.assert_allclose(L @ L.T, C, atol=1e-12)


# ------------------------------------------------------ matvec and CG ----

@pytest.mark.parametrize("dim,cells,kind,corr,aniso,matrix", [
    (2, [6, 5], "exponential", [0.2], "none", "dct"),
    (2, [5, 4], "gaussian", [0.1, 0.02, 0.0, 0.08], "geometric", "r2c"),
    (3, [3, 4, 3], "matern32", [0.2, 0.1, 0.1], "axiparallel", "dct"),
])
def test_matvec_is_dense_covariance_over_next(dim, cells, kind, corr, aniso, matrix):
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, matrix=matrix,
             approximate=True)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0] * dim, kind, aniso, corr)
    x = np.random.default_rng(3).standard_normal(C.shape[0])
    np.testing.assert_allclose(orc.matvec(d, x), C @ x / s.extended_size, atol=1e-14)


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,tr", [
    (2, [5, 4], "exponential", [0.2], "none", -1),
    (2, [4, 6], "gaussian", [0.2, 0.1], "axiparallel", 0),
    (3, [3, 4, 2], "matern32", [0.2, 0.1, 0.15], "axiparallel", -1),
    (3, [4, 2, 3], "exponential", [0.3], "none", 0),
])
def test_dctdst_matvec_route(dim, cells, kind, corr, aniso, tr):
    """The DCT/DST route of multiplyExtended (per parity type: setComponent,
    r2r forward / N_ext, times f(lambda), r2r backward, added into the output)
    sums to R f(C_ext) E x / N_ext: the multiply equals the dense covariance
    C x / N_ext, and every operator equals the R2C-field route on the same
    DCT eigenvalues (independent transforms)."""
    kw = dict(covariance=kind, corr_length=corr, anisotropy=aniso, matrix="dct",
              approximate=True, transposed=tr)
    d = desc(dim, cells, field="dctdst", **kw)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0] * dim, kind, aniso, corr)
    x = np.random.default_rng(7).standard_normal(C.shape[0])
    np.testing.assert_allclose(orc.matvec(d, x), C @ x / s.extended_size, atol=1e-14)
    d2 = desc(dim, cells, field="r2c", **dict(kw, transposed=-1 if dim == 3 else 0))
    for op in ("multiply", "root", "inverse_embed"):
        want = orc.matvec(d2, x, op)
        np.testing.assert_allclose(orc.matvec(d, x, op), want,
                                   atol=1e-12 * np.abs(want).max(), err_msg=op)


def test_root_and_inverse_are_exact_when_periodic():
    d = desc(2, [8, 6], covariance="exponential", corr_length=0.1, periodic=True)
    x = np.random.default_rng(4).standard_normal(48)
    y = orc.matvec(d, x)
    np.testing.assert_allclose(orc.matvec(d, orc.matvec(d, x, "root"), "root"), y, atol=1e-13)
    np.testing.assert_allclose(orc.matvec(d, y, "inverse_embed"), x, atol=1e-10)


def test_multiply_inverse_solves_dense_system():
    cells = [6, 6]
    d = desc(2, cells, covariance="exponential", corr_length=0.2)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0, 1.0], "exponential", "none", [0.2]) / s.extended_size
    b = np.random.default_rng(5).standard_normal(36)
    x, it = orc.multiply_inverse(d, b)
    want = np.linalg.solve(C, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-6
    assert 0 < it <= 100
    # zero right-hand side returns the input unchanged (matrix.hh:193-196)
    z = np.full(36, 1e-11)
    x0, it0 = orc.multiply_inverse(d, z)
    assert it0 == 0 and np.array_equal(x0, z)


@pytest.mark.parametrize("mode", ["plain", "embed_pcg"])
def test_pcg_solves_dense_system(mode):
    cells = [6, 5, 4]
    d = desc(3, cells, covariance="exponential", corr_length=0.15, nugget=1e-3)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0] * 3, "exponential", "none", [0.15])
    C = (C + 1e-3 * np.eye(C.shape[0])) / s.extended_size
    b = np.random.default_rng(6).standard_normal(C.shape[0])
    x, it, rr = orc.pcg(d, b, mode=mode, maxit=500, rel_tol=1e-10)
    assert rr <= 1e-10
    want = np.linalg.solve(C, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-7


# --------------------------------------------------------- DFT backends --
# DFTMatrixBackend (dftmatrixbackend.hh) and DFTFieldBackend
# (dftfieldbackend.hh): the 1-D defaults, selectable in 2D/3D.

def _dft_storage_k(s, N, tr):
    """Frequency (k0, k1, k2) of storage index s: x fastest, the two last
    axes swapped when transposed (transposeIfNeeded)."""
    dim = len(N)
    if dim == 1:
        return (s,)
    if dim == 2:
        return (s // N[1], s % N[1]) if tr else (s % N[0], s // N[0])
    if tr:
        return (s % N[0], s // (N[0] * N[2]), (s // N[0]) % N[2])
    return (s % N[0], (s // N[0]) % N[1], s // (N[0] * N[1]))


def _dft_store(full, tr):
    """numpy (k_{d-1}, ..., k0) array -> the reference storage order."""
    return full.swapaxes(0, 1) if (tr and full.ndim > 1) else full


DFT_CASES = [
    (1, [12], "exponential", [0.2], "none", 0),
    (1, [10], "spherical", [0.3], "geometric", 0),
    (2, [5, 4], "gaussian", [0.1, 0.02, 0.0, 0.08], "geometric", 1),
    (2, [4, 6], "exponential", [0.3], "none", 0),
    (3, [3, 2, 4], "matern32", [0.2, 0.1, 0.15], "axiparallel", 1),
    (3, [2, 3, 2], "exponential", [0.25], "none", 0),
]


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,tr", DFT_CASES)
def test_eigs_dft_vs_library(dim, cells, kind, corr, aniso, tr):
    """Every entry of the grid: the real part of numpy's fftn of the wrapped
    covariance / N_ext, in the reference storage order, classified and
    clamped over all N_ext entries (matrix.hh:266-291)."""
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, matrix="dft",
             field="dft", transposed=tr, approximate=True)
    v, st, _ = orc.eigs(d, allow_negative=True)
    s = orc.sizes(d)
    assert s.eig_entries == s.extended_size == v.size
    assert s.transposed == (tr and dim > 1)
    full = nh.full_eigs(cells, [1.0] * dim, kind, aniso, corr).real
    flat = np.ascontiguousarray(_dft_store(full, tr)).ravel()
    thr = 1e-14
    assert st["negative"] == int(np.sum(flat <= -thr))
    assert st["zero"] == int(np.sum(np.abs(flat) < thr))
    assert st["small"] == int(np.sum((flat >= thr) & (flat < 1e-6)))
    np.testing.assert_allclose(v, np.maximum(flat, 0.0), atol=1e-13 * np.abs(flat).max())


def _circulant_cov(cells, kind, aniso, corr):
    """Covariance of the embedded model on the cells from numpy's spectrum
    (clamped at 0): C[i, j] = sum_k lambda_k e^{2 pi i k (x_i - x_j) / N}."""
    full = np.maximum(nh.full_eigs(cells, [1.0] * len(cells), kind, aniso, corr).real, 0.0)
    c = np.fft.ifftn(full).real * full.size
    x = np.stack(np.meshgrid(*[np.arange(n) for n in cells[::-1]], indexing="ij")[::-1],
                 axis=-1).reshape(-1, len(cells))
    N = np.array([2 * n for n in cells])
    dx = (x[:, None, :] - x[None, :, :]) % N
    return c[tuple(dx[..., i] for i in range(len(cells) - 1, -1, -1))]


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,tr", DFT_CASES)
def test_generate_dft_covariance_identity(dim, cells, kind, corr, aniso, tr):
    """Field and spare field are each distributed with the embedded covariance
    and uncorrelated with each other: L_re L_re^T = L_im L_im^T = C,
    L_re L_im^T = 0 (lambda (xi_1 + i xi_2) through the backward transform)."""
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, matrix="dft",
             field="dft", transposed=tr, approximate=True)
    n = int(orc.sizes(d).noise_per_field)
    assert n == 2 * int(np.prod([2 * c for c in cells]))
    cols = []
    for j in range(n):
        z = np.zeros(n)
        z[j] = 1.0
        cols.append(orc.generate(d, z))
    Lre = np.stack([c[0] for c in cols], axis=1)
    Lim = np.stack([c[1] for c in cols], axis=1)
    C = _circulant_cov(cells, kind, aniso, corr)
    np.testing.assert_allclose(Lre @ Lre.T, C, atol=1e-12)
    np.testing.assert_allclose(Lim @ Lim.T, C, atol=1e-12)
    np.testing.assert_allclose(Lre @ Lim.T, 0.0, atol=1e-12)


@pytest.mark.parametrize("dim,cells,tr", [(1, [10], 0), (2, [4, 3], 0), (2, [4, 3], 1),
                                          (3, [2, 3, 4], 0), (3, [2, 3, 4], 1)])
def test_generate_dft_single_mode_noise_order(dim, cells, tr):
    """A unit sample in slot 0 (1) of storage entry s gives the field
    sqrt(l) cos(ph) (-sqrt(l) sin(ph)) and the spare field sqrt(l) sin(ph)
    (sqrt(l) cos(ph)), ph = 2 pi k.x / N, with k the storage order of the
    reference (transposeIfNeeded swaps the two last axes)."""
    d = desc(dim, cells, covariance="exponential", corr_length=0.3, matrix="dft", field="dft",
             transposed=tr)
    N = [2 * c for c in cells]
    n = int(orc.sizes(d).noise_per_field)
    full = nh.full_eigs(cells, [1.0] * dim, "exponential", "none", [0.3]).real
    x = np.stack(np.meshgrid(*[np.arange(c) for c in cells[::-1]], indexing="ij")[::-1],
                 axis=-1).reshape(-1, dim)
    rng = np.random.default_rng(8)
    for s in sorted(set([0, 1, n // 2 - 1] + list(rng.integers(0, n // 2, 6)))):
        k = _dft_storage_k(int(s), N, tr)
        lam = full[tuple(k[::-1])]
        ph = 2 * np.pi * sum(k[i] * x[:, i] / N[i] for i in range(dim))
        for slot in (0, 1):
            z = np.zeros(n)
            z[2 * s + slot] = 1.0
            f, spare = orc.generate(d, z)
            a = np.sqrt(lam)
            want_f = a * np.cos(ph) if slot == 0 else -a * np.sin(ph)
            want_s = a * np.sin(ph) if slot == 0 else a * np.cos(ph)
            np.testing.assert_allclose(f, want_f, atol=1e-13)
            np.testing.assert_allclose(spare, want_s, atol=1e-13)


@pytest.mark.parametrize("dim,cells,kind,corr,aniso,tr", [c for c in DFT_CASES
                                                          if c[2] != "gaussian"])
def test_dft_matvec_is_dense_covariance_over_next(dim, cells, kind, corr, aniso, tr):
    d = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso, matrix="dft",
             field="dft", transposed=tr)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0] * dim, kind, aniso, corr)
    x = np.random.default_rng(3).standard_normal(C.shape[0])
    np.testing.assert_allclose(orc.matvec(d, x), C @ x / s.extended_size, atol=1e-14)
    if dim > 1:
        # the complex route of the DFT field equals the R2C route on the
        # same (DFT) matrix
        r = desc(dim, cells, covariance=kind, corr_length=corr, anisotropy=aniso,
                 matrix="dft", field="r2c")
        for op in ("multiply", "root", "inverse_embed"):
            np.testing.assert_allclose(orc.matvec(d, x, op), orc.matvec(r, x, op),
                                       rtol=1e-12, atol=1e-12)


def test_dft_1d_root_and_inverse_when_periodic_and_cg():
    d = desc(1, [64], covariance="exponential", corr_length=0.1, periodic=True)
    assert d.matrix_backend == pfb.MATRIX_BACKENDS["dft"]
    x = np.random.default_rng(4).standard_normal(64)
    y = orc.matvec(d, x)
    np.testing.assert_allclose(orc.matvec(d, orc.matvec(d, x, "root"), "root"), y, atol=1e-13)
    np.testing.assert_allclose(orc.matvec(d, y, "inverse_embed"), x, atol=1e-10)
    cells = [40]
    d = desc(1, cells, covariance="exponential", corr_length=0.2)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0], "exponential", "none", [0.2]) / s.extended_size
    b = np.random.default_rng(5).standard_normal(40)
    xs, it = orc.multiply_inverse(d, b)
    want = np.linalg.solve(C, b)
    assert np.linalg.norm(xs - want) / np.linalg.norm(want) < 1e-6


def test_dft_backend_rules():
    with pytest.raises(RuntimeError):
        orc.sizes(desc(1, [8], matrix="dct", field="dft"))
    with pytest.raises(RuntimeError):
        orc.sizes(desc(1, [8], matrix="dft", field="r2c"))


# ----------------------------------------------------- value transforms --

def test_value_transforms():
    v = np.linspace(-2, 2, 41)
    np.testing.assert_array_equal(orc.value_transform(v, "none"), v)
    np.testing.assert_allclose(orc.value_transform(v, "logNormal"), np.exp(v), rtol=1e-15)
    np.testing.assert_array_equal(orc.value_transform(v, "foldedNormal"), np.abs(v))
    np.testing.assert_array_equal(orc.value_transform(v, "sign"), np.where(v < 0, -1.0, 1.0))
    np.testing.assert_array_equal(orc.value_transform(v, "sign", 0.5),
                                  np.where(v < 0.5, -1.0, 1.0))
    lam = 0.3
    np.testing.assert_allclose(orc.value_transform(v, "boxCox", lam),
                               np.power(v * lam + 1, -lam), rtol=1e-14)


# ------------------------------------------ learned preconditioner (oracle) --

def test_mlp_backward_matches_central_differences():
    dims = [6, 16, 16, 16, 8]
    rng = np.random.default_rng(3)
    n = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    p = rng.normal(0, 0.4, n)
    x = rng.normal(size=(7, 6))
    y = rng.normal(size=(7, 8))
    loss, g, _ = orc.mlp(dims, p, x, y)
    h = 1e-6
    for i in rng.choice(n, 60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp(dims, pp, x, y, grads=False)[0] - orc.mlp(dims, pm, x, y, grads=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # forward = explicit layer composition with exact GELU
    a = x
    off = 0
    for l in range(len(dims) - 1):
        W = p[off:off + dims[l] * dims[l + 1]].reshape(dims[l + 1], dims[l])
        off += dims[l] * dims[l + 1]
        b = p[off:off + dims[l + 1]]
        off += dims[l + 1]
        z = a @ W.T + b
        a = 0.5 * z * (1 + special.erf(z / np.sqrt(2))) if l + 2 < len(dims) else z
    assert abs(loss - np.mean((a - y) ** 2)) < 1e-12


def test_spectral_targets_vs_numpy():
    cells = [8, 6, 4]
    d = desc(3, cells, covariance="exponential", corr_length=0.2, nugget=1e-3)
    s = orc.sizes(d)
    floor = 1e-3 / s.extended_size
    got = orc.spectral_targets(d, 12, floor)
    v, _, _ = orc.eigs(d)
    P = [c + 1 for c in cells]
    lam = v.reshape(P[1], P[2], P[0]).transpose(1, 0, 2)  # natural [kz][ky][kx]
    kz, ky, kx = np.meshgrid(*[np.arange(q) for q in P[::-1]], indexing="ij")
    r = np.sqrt(((kx / cells[0]) ** 2 + (ky / cells[1]) ** 2 + (kz / cells[2]) ** 2) / 3)
    b = np.minimum((r * 12).astype(int), 11)
    want = np.array([-np.log(np.maximum(lam[b == i], floor).mean()) for i in range(12)])
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_learned_pcg_with_method_targets_solves_dense_system():
    cells = [6, 5, 4]
    d = desc(3, cells, covariance="exponential", corr_length=0.15, nugget=1e-3)
    s = orc.sizes(d)
    C = nh.dense_cov(cells, [1.0] * 3, "exponential", "none", [0.15])
    C = (C + 1e-3 * np.eye(C.shape[0])) / s.extended_size
    b = np.random.default_rng(9).standard_normal(C.shape[0])
    t = orc.spectral_targets(d, 16, 1e-3 / s.extended_size)
    x, it, rr = orc.pcg_learned(d, b, t, maxit=500, rel_tol=1e-10)
    assert rr <= 1e-10
    want = np.linalg.solve(C, b)
    assert np.linalg.norm(x - want) / np.linalg.norm(want) < 1e-7
    _, it_plain, _ = orc.pcg(d, b, mode="plain", maxit=500, rel_tol=1e-10)
    assert it < it_plain


# --- merge embedding (matrix.hh:652-775, covariance.hh:147-290) -------------

@pytest.mark.parametrize("k", range(7))
def test_smoothstep_sigmoid_is_regularized_incomplete_beta(k):
    """The smoothstep taper with k recursions is I_t(k+1, k+1) of the clamped
    value t (the Hermite interpolation matching k derivatives: k = 1 gives
    3t^2 - 2t^3, k = 2 gives 6t^5 - 15t^4 + 10t^3)."""
    one, zero = 0.3, 1.7
    for x in np.linspace(0.0, 2.0, 41):
        t = 1.0 if abs(x) < one else 0.0 if abs(x) > zero else (zero - abs(x)) / (zero - one)
        assert orc.sigmoid("smoothstep", one, zero, x, k) == pytest.approx(
            special.betainc(k + 1, k + 1, t), abs=1e-14)
        assert orc.sigmoid("smoothstep", one, zero, -x, k) == orc.sigmoid(
            "smoothstep", one, zero, x, k)
    assert orc.sigmoid("clamp", one, zero, 1.0) == pytest.approx(0.5)


def test_smooth_sigmoid_shape():
    """C-infinity taper: 1 below the one edge, 0 beyond the zero edge,
    s(x) + s(one + zero - x) = 1 in between (so 1/2 at the midpoint),
    decreasing."""
    one, zero = 0.5, 2.0
    assert orc.sigmoid("smooth", one, zero, 0.2) == 1.0
    assert orc.sigmoid("smooth", one, zero, 2.5) == 0.0
    xs = np.linspace(one, zero, 31)
    v = np.array([orc.sigmoid("smooth", one, zero, x) for x in xs])
    w = np.array([orc.sigmoid("smooth", one, zero, one + zero - x) for x in xs])
    np.testing.assert_allclose(v + w, 1.0, atol=1e-15)
    assert orc.sigmoid("smooth", one, zero, 1.25) == pytest.approx(0.5, abs=1e-15)
    assert np.all(np.diff(v) <= 0)


MERGE_CASES = [
    (2, [12, 10], dict(covariance="matern52", corr_length=0.25, factor=3, sigmoid_combine="tensor")),
    (2, [12, 10], dict(covariance="exponential", corr_length=0.3, factor=4, sigmoid="smooth")),
    (3, [6, 5, 4], dict(covariance="matern32", corr_length=[0.3, 0.2, 0.25],
                        anisotropy="axiparallel", factor=3)),
    (2, [10, 8], dict(covariance="gaussian", corr_length=[0.2, 0.1, 0.05, 0.3],
                      anisotropy="geometric", factor=3, merge_recursions=3,
                      sigmoid_combine="radial")),
]


@pytest.mark.parametrize("dim,cells,kw", MERGE_CASES)
@pytest.mark.parametrize("matrix", ["dct", "r2c"])
def test_merge_keeps_the_covariance_on_the_domain(dim, cells, kw, matrix):
    """Inside the original domain (every x_i < ext_i; for a radial taper also
    |x / ext| < sigmoidStart sqrt(d)) the merged entry is the covariance
    itself: the unmirrored copy has weight 1 and every mirror image weight 0
    (matrix.hh:652-775)."""
    d = desc(dim, cells, matrix=matrix, approximate=True, periodization="merge", **kw)
    c = desc(dim, cells, matrix=matrix, approximate=True, **kw)
    gm, gc = orc.cov_grid(d), orc.cov_grid(c)
    idx = np.stack(np.meshgrid(*[np.arange(s) for s in gm.shape], indexing="ij"), -1)  # [z][y][x]
    pos = idx[..., ::-1].astype(float) / np.array(cells, float)  # x / ext, x first
    inside = np.all(pos < 1.0, axis=-1)
    if kw.get("sigmoid_combine", "radialPlus") != "tensor":
        inside &= np.sqrt((pos ** 2).sum(-1)) < np.sqrt(dim) - 1e-9
    assert inside.sum() > 0
    np.testing.assert_allclose(gm[inside], gc[inside], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("dim,cells,kw", MERGE_CASES)
def test_merge_grid_is_mirror_symmetric(dim, cells, kw):
    """The merged R2C grid is even about the centre of every axis
    (G[.., i] = G[.., N - i]): the mirror images of x and of N h - x are the
    same set."""
    d = desc(dim, cells, matrix="r2c", approximate=True, periodization="merge", **kw)
    g = orc.cov_grid(d)
    for ax in range(dim):
        a = g.shape[dim - 1 - ax]  # axis ax of x-fastest data sits at numpy axis dim-1-ax
        flipped = np.take(g, [(-i) % a for i in range(a)], axis=dim - 1 - ax)
        np.testing.assert_allclose(g, flipped, rtol=1e-13, atol=1e-15)


def test_merge_with_factor_two_is_classical_on_the_dct_grid():
    """With embedding factor 2 the taper interval (sigmoidStart 1,
    sigmoidEnd factor - 1 = 1) has no width: on the DCT half grid the merge
    entries are the classical ones."""
    kw = dict(covariance="matern52", corr_length=0.3, approximate=True)
    gm = orc.cov_grid(desc(2, [16, 12], periodization="merge", sigmoid_combine="tensor", **kw))
    gc = orc.cov_grid(desc(2, [16, 12], **kw))
    np.testing.assert_allclose(gm, gc, rtol=1e-13, atol=1e-15)


def test_merge_removes_the_negative_eigenvalues_of_mirroring():
    """The purpose of the smooth embedding: Matern-5/2 with l = 0.25 on a
    factor-3 embedding has negative eigenvalues with classical mirroring and
    none with the merge embedding."""
    kw = dict(covariance="matern52", corr_length=0.25, factor=3, approximate=True)
    _, st_c, _ = orc.eigs(desc(2, [64, 64], **kw), allow_negative=True)
    _, st_m, status = orc.eigs(desc(2, [64, 64], periodization="merge",
                                    sigmoid_combine="tensor", **kw), allow_negative=True)
    assert st_c["negative"] > 0
    assert st_m["negative"] == 0 and status == 0 and st_m["smallest"] > 0


def test_merge_default_recursions_rule():
    """embedding.mergeRecursions 99 (the default) has a count for the Matern
    families, exponential, cauchy and their generalizations only; others are
    rejected like the reference does (matrix.hh:679-705)."""
    with pytest.raises(RuntimeError):
        orc.cov_grid(desc(2, [8, 8], covariance="gaussian", factor=3, periodization="merge"))
    orc.cov_grid(desc(2, [8, 8], covariance="gaussian", factor=3, periodization="merge",
                      merge_recursions=2))
    # explicit counts 1 and 2 through the taper change the entries beyond the domain
    g1 = orc.cov_grid(desc(2, [8, 8], covariance="exponential", corr_length=0.5, factor=3,
                           matrix="r2c", periodization="merge", merge_recursions=1))
    g2 = orc.cov_grid(desc(2, [8, 8], covariance="exponential", corr_length=0.5, factor=3,
                           matrix="r2c", periodization="merge", merge_recursions=2))
    assert np.abs(g1 - g2).max() > 1e-6


# Fold embedding (matrix.hh:790-888): beyond extension-scaled norm 1 the
# coordinates move towards the origin, to norm 1 + the integral of the taper
# sigmoid(1, beta, x) over [1, norm], beta = foldMaxFactor * factor / (2 sqrt(dim)).

@pytest.mark.parametrize("sigmoid,rec", [("smoothstep", 0), ("smoothstep", 1), ("smoothstep", 4),
                                         ("smoothstep", 9), ("smooth", 0)])
@pytest.mark.parametrize("beta", [1.3, 2.0, 3.5])
def test_fold_integral_matches_quadrature(sigmoid, rec, beta):
    """The oracle's integral (closed-form antiderivative for smoothstep,
    adaptive Simpson for smooth) against scipy's QUADPACK on the sigmoid."""
    from scipy import integrate
    kind = "smoothstep" if sigmoid == "smoothstep" else "smooth"
    for norm in (1.0 + 1e-3, 1.0 + 0.37 * (beta - 1.0), 1.0 + 0.8 * (beta - 1.0)):
        want, _ = integrate.quad(lambda x: orc.sigmoid(kind, 1.0, beta, x, rec), 1.0, norm,
                                 epsabs=1e-15, epsrel=1e-13, limit=200)
        got = orc.fold_integral(sigmoid, beta, norm, rec)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (norm, got, want)


@pytest.mark.parametrize("sigmoid,rec", [("smoothstep", 0), ("smoothstep", 3), ("smoothstep", 6),
                                         ("smooth", 0)])
def test_fold_integral_closed_forms(sigmoid, rec):
    """Both tapers satisfy s(x) + s(1 + beta - x) = 1, so the full integral is
    (beta - 1) / 2, and it stays there beyond beta; no recursion: the taper is
    linear, integral (beta - 1)(t - t^2 / 2) at t = (norm - 1) / (beta - 1);
    beta <= 1 or norm <= 1: zero."""
    for beta in (1.25, 2.0, 4.0):
        for norm in (beta, beta + 0.5, 10.0):
            assert abs(orc.fold_integral(sigmoid, beta, norm, rec) - 0.5 * (beta - 1.0)) < 1e-12
        if sigmoid == "smoothstep" and rec == 0:
            for t in (0.1, 0.5, 0.9):
                want = (beta - 1.0) * (t - 0.5 * t * t)
                got = orc.fold_integral(sigmoid, beta, 1.0 + t * (beta - 1.0), 0)
                assert abs(got - want) < 1e-13
    assert orc.fold_integral(sigmoid, 0.7, 3.0, rec) == 0.0
    assert orc.fold_integral(sigmoid, 1.0, 3.0, rec) == 0.0
    assert orc.fold_integral(sigmoid, 2.0, 1.0, rec) == 0.0


@pytest.mark.parametrize("sigmoid", ["smoothstep", "smooth"])
def test_fold_grid_invariants(sigmoid):
    """Exponential covariance, 2D, equal extensions, factor 4 (beta = sqrt 2),
    DCT half grid: entries with norm <= 1 are the classical ones; entries with
    norm >= beta all take C at radius sqrt(dim) (1 + (beta - 1) / 2);
    in between, C at the coordinates scaled to 1 + integral."""
    cells, ext, l = [24, 20], [1.0, 1.0], 0.3
    kw = dict(covariance="exponential", corr_length=l, factor=4, matrix="dct", extensions=ext)
    fold = orc.cov_grid(desc(2, cells, periodization="fold", sigmoid=sigmoid, **kw))
    classic = orc.cov_grid(desc(2, cells, **kw))
    beta = 4 / (2 * np.sqrt(2.0))
    rec = 2  # exponential: default recursion count
    h = [ext[i] / cells[i] for i in range(2)]
    inner = beyond = between = 0
    for j in range(fold.shape[0]):
        for i in range(fold.shape[1]):
            x, y = i * h[0], j * h[1]
            norm = np.sqrt((x * x / ext[0] ** 2 + y * y / ext[1] ** 2) / 2)
            if norm <= 1:
                assert fold[j, i] == classic[j, i]
                inner += 1
                continue
            if norm >= beta:
                r = np.sqrt(2.0) * (1 + 0.5 * (beta - 1))
                beyond += 1
            else:
                r = np.hypot(x, y) * (1 + orc.fold_integral(sigmoid, beta, norm, rec)) / norm
                between += 1
            assert abs(fold[j, i] - np.exp(-r / l)) < 1e-14, (i, j)
    assert inner and beyond and between


def test_fold_rejects_unknown_periodization_and_family_without_default_count():
    d = desc(2, [8, 8], covariance="exponential")
    d.periodization = 4
    with pytest.raises(RuntimeError):
        orc.cov_grid(d)
    with pytest.raises(RuntimeError):
        orc.cov_grid(desc(2, [8, 8], covariance="gaussian", factor=3, periodization="fold"))
    orc.cov_grid(desc(2, [8, 8], covariance="gaussian", factor=3, periodization="fold",
                      merge_recursions=3))


# Cofold (matrix.hh:905-1013): the classical entries beyond the domain (max-norm
# > 1) are lifted smoothly: below alpha (largest border entry) to the offset,
# in [alpha, beta) along the integral of the codomain taper, beta =
# min(cofoldMaxFactor * alpha, smallest entry on the domain).

def _cofold_expect(classic, cells, ext, factor, sigmoid, rec, max_factor):
    """Independent restatement on the DCT half grid with scipy quadrature."""
    from scipy import integrate
    N = [c * factor for c in cells]
    h = [ext[i] / cells[i] for i in range(2)]
    j, i = np.meshgrid(np.arange(classic.shape[0]), np.arange(classic.shape[1]), indexing="ij")
    x, y = i * h[0], j * h[1]
    norm = np.maximum(np.abs(x) / ext[0], np.abs(y) / ext[1])
    border = ((0.5 * ext[0] * factor - np.abs(x)) < ext[0] / N[0]) | \
             ((0.5 * ext[1] * factor - np.abs(y)) < ext[1] / N[1])
    alpha = classic[(norm > 1) & border].max()
    beta = min(max_factor * alpha, classic[norm <= 1].min())
    w = beta - alpha
    offset = 0.5 * (alpha + beta)  # the taper's symmetry: its integral over [alpha, beta] is w / 2
    want = classic.copy()
    for (jj, ii) in zip(*np.nonzero(norm > 1)):
        v = classic[jj, ii]
        if v < alpha:
            want[jj, ii] = offset
        elif v < beta:
            I, _ = integrate.quad(lambda t: orc.sigmoid(sigmoid, 0.0, w, beta - t, rec), alpha, v,
                                  epsabs=1e-17, epsrel=1e-13)
            want[jj, ii] = I + offset
    return want, norm, alpha, beta


@pytest.mark.parametrize("sigmoid,rec,max_factor", [("smoothstep", 1, 5.0), ("smoothstep", 4, 1.2),
                                                    ("smooth", 0, 5.0)])
def test_cofold_grid_matches_restatement(sigmoid, rec, max_factor):
    cells, ext, factor = [24, 20], [1.0, 1.0], 3
    kw = dict(covariance="exponential", corr_length=0.3, factor=factor, matrix="dct",
              extensions=ext)
    classic = orc.cov_grid(desc(2, cells, **kw))
    got = orc.cov_grid(desc(2, cells, periodization="cofold", sigmoid=sigmoid,
                            merge_recursions=rec if rec else -1, fold_max_factor=max_factor,
                            **kw))
    want, norm, alpha, beta = _cofold_expect(classic, cells, ext, factor, sigmoid, rec, max_factor)
    assert np.array_equal(got[norm <= 1], classic[norm <= 1])
    assert np.abs(got - want).max() <= 1e-13 * np.abs(classic).max()
    # the lifted entries lie in [offset, beta) and keep the order of the originals
    outer = norm > 1
    lifted = outer & (classic < beta)
    assert lifted.any() and (classic[outer] < alpha).any()
    assert got[lifted].min() >= 0.5 * (alpha + beta) - 1e-15
    order = np.argsort(classic[lifted], kind="stable")
    assert np.all(np.diff(got[lifted][order]) >= -1e-16)


def test_taper_integral_matches_quadrature():
    from scipy import integrate
    for sigmoid, rec in (("smoothstep", 1), ("smoothstep", 5), ("smooth", 0)):
        for one, zero in ((0.0, 2e-3), (0.5, 0.9)):
            for frac in (0.2, 0.6, 1.0):
                hi = one + frac * (zero - one)
                want, _ = integrate.quad(lambda x: orc.sigmoid(sigmoid, one, zero, x, rec), one,
                                         hi, epsabs=1e-18, epsrel=1e-13)
                got = orc.taper_integral(sigmoid, one, zero, hi, rec)
                assert abs(got - want) <= 1e-12 * abs(want), (sigmoid, rec, one, zero, frac)
