"""numpy/scipy helpers that build independent checks for the oracle and the
HIP path: textbook covariance closed forms, the wrapped embedding grid, dense
covariance matrices (brute force), library FFT/DCT routines."""
from __future__ import annotations

import numpy as np
from scipy import special


def textbook_cov(kind: str, h: np.ndarray, var: float = 1.0, dim: int = 2, nu: float = 1.0,
                 gamma: float = 1.0, alpha: float = 1.0, beta: float = 1.0) -> np.ndarray:
    """Covariance families as functions of the scaled distance h (standard
    textbook definitions: exponential, Gaussian, Matern(nu) with
    K_nu from scipy, spherical (1D/2D/3D), etc.)."""
    h = np.asarray(h, dtype=np.float64)
    if kind == "exponential":
        return var * np.exp(-h)
    if kind == "gaussian":
        return var * np.exp(-h ** 2)
    if kind == "matern32":
        return var * (1 + np.sqrt(3) * h) * np.exp(-np.sqrt(3) * h)
    if kind == "matern52":
        return var * (1 + np.sqrt(5) * h + 5 * h ** 2 / 3) * np.exp(-np.sqrt(5) * h)
    if kind == "matern":
        z = np.sqrt(2 * nu) * np.maximum(h, 1e-300)
        v = var * 2 ** (1 - nu) / special.gamma(nu) * z ** nu * special.kv(nu, z)
        return np.where(h < 1e-10, var, v)
    if kind == "spherical":
        hc = np.minimum(h, 1.0)
        if dim == 3:
            v = var * (1 - 1.5 * hc + 0.5 * hc ** 3)
        elif dim == 2:
            v = var * (1 - 2 / np.pi * (hc * np.sqrt(1 - hc ** 2) + np.arcsin(hc)))
        else:
            v = var * (1 - hc)  # 1-D spherical (triangular) model
        return np.where(h > 1, 0.0, v)
    if kind == "cauchy":
        return var * (1 + h ** 2) ** -3
    if kind == "generalizedCauchy":
        return var * (1 + h ** alpha) ** (-beta)
    if kind == "gammaExponential":
        return var * np.exp(-h ** gamma)
    if kind == "dampedOscillation":
        return var * np.exp(-h) * np.cos(h / np.sqrt(3) if dim == 3 else h)
    if kind == "cubic":
        v = var * (1 - 7 * h ** 2 + 8.75 * h ** 3 - 3.5 * h ** 5 + 0.75 * h ** 7)
        return np.where(h > 1, 0.0, v)
    raise ValueError(kind)


def scale_coords(x: np.ndarray, anisotropy: str, corr) -> np.ndarray:
    """Map coordinates (..., dim) to the scaled space: x/l, x_i/l_i or A^{-1} x."""
    dim = x.shape[-1]
    corr = np.atleast_1d(np.asarray(corr, dtype=np.float64))
    if anisotropy == "none":
        return x / corr[0]
    if anisotropy == "axiparallel":
        return x / corr[:dim]
    A = corr[: dim * dim].reshape(dim, dim)
    return x @ np.linalg.inv(A).T


def cov_at(x: np.ndarray, kind: str, anisotropy: str, corr, **kw) -> np.ndarray:
    if kind == "separableExponential":
        t = scale_coords(x, anisotropy, corr)
        return kw.get("var", 1.0) * np.exp(-np.abs(t).sum(-1))
    t = scale_coords(x, anisotropy, corr)
    return textbook_cov(kind, np.sqrt((t ** 2).sum(-1)), dim=x.shape[-1], **kw)


def cell_coords(cells, ext) -> np.ndarray:
    """Lower-left corners of all cells, x fastest: shape (prod n, dim)."""
    dim = len(cells)
    axes = [np.arange(cells[i]) * (ext[i] / cells[i]) for i in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")
    return np.stack([g.ravel() for g in grids[::-1]], axis=-1)


def dense_cov(cells, ext, kind, anisotropy, corr, **kw) -> np.ndarray:
    """Brute-force covariance matrix C[i, j] = C(x_i - x_j) on the original grid."""
    x = cell_coords(cells, ext)
    d = x[:, None, :] - x[None, :, :]
    return cov_at(d.reshape(-1, x.shape[1]), kind, anisotropy, corr, **kw).reshape(
        x.shape[0], x.shape[0])


def wrapped_grid(cells, ext, factor=2):
    """Coordinates of the full extended grid with the classical wrap rule:
    coordinate > half the extended length is shifted by the extended length.
    Shape (N_{d-1}, ..., N_0, dim), x fastest (numpy last axis)."""
    dim = len(cells)
    axes = []
    for i in range(dim):
        N = factor * cells[i]
        h = ext[i] / cells[i]
        c = np.arange(N) * h
        c = np.where(c > 0.5 * ext[i] * factor, c - ext[i] * factor, c)
        axes.append(c)
    grids = np.meshgrid(*axes[::-1], indexing="ij")
    return np.stack(grids[::-1], axis=-1)


def full_eigs(cells, ext, kind, anisotropy, corr, factor=2, **kw) -> np.ndarray:
    """All N_ext eigenvalues of the circulant embedding divided by N_ext,
    numpy array shaped (N_{d-1}, ..., N_0): fftn of the wrapped covariance."""
    g = wrapped_grid(cells, ext, factor)
    c = cov_at(g.reshape(-1, len(cells)), kind, anisotropy, corr, **kw).reshape(g.shape[:-1])
    return np.fft.fftn(c) / c.size


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
