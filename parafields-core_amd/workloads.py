"""The synthetic workloads the benchmark measures and the GPU tests check at
full size, in one place: the 3D ensemble (Matern-3/2 on a 512^3 embedding),
the uneven batch of 2D fields, the covariance-inverse case and the 1024^3
DCT/DST field. Each returns descriptors for pf_binding."""
from __future__ import annotations

import numpy as np

import pf_binding as pfb

# 3D ensemble: 256^3 cells on [0,1]^3, Matern nu = 1.5, variance 2,
# axiparallel lengths (0.1, 0.1, 0.05), factor 2 -> 512^3, DCT matrix + R2C
# field; its eigenvalues have small negative entries (embedding.approximate).
ENSEMBLE_CELLS = [256, 256, 256]
ENSEMBLE_KW = dict(covariance="matern32", variance=2.0, anisotropy="axiparallel",
                   corr_length=[0.1, 0.1, 0.05], approximate=True)
ENSEMBLE_SEED = 7

UNEVEN_SIDES = [64, 96, 128, 192, 256, 384, 512, 768, 1024]


def uneven_batch(nfields: int = 1000, seed: int = 3, sides=UNEVEN_SIDES, precision="fp32"):
    """2D fields on [0,1]^2 with sides drawn independently and uniformly from
    `sides` (log-uniform over 2^k and 3*2^k), Gaussian covariance with full
    geometric anisotropy A = R(theta) diag(L, L/r), theta ~ U[0, pi),
    r ~ U[1, 8], L ~ U[0.02, 0.2]; R2C matrix with the full real-part check
    (the reference's interleaved check would turn Gaussian fields into NaN).
    Returns (descs, shapes, algorithmic bytes per field).

    Algorithmic bytes of a field (embedding N0 = 2 n0, N1 = 2 n1, H = n0 + 1,
    h = n1 + 1 rows of the even part): covariance rows written (8 N0 h),
    read by the r2c along x (8 N0 h) which writes the half spectrum
    (16 H h); the real y pass reads it (16 H h) and writes the FP32 root
    table (4 H N1); the generation pass reads the table (4 H N1) and writes
    the y-pruned FP32 spectrum (8 H n1); the x c2r reads it (8 H n1) and
    writes the field (4 n0 n1)."""
    rng = np.random.default_rng(seed)
    descs, shapes, falg = [], [], []
    for _ in range(nfields):
        n0, n1 = (int(rng.choice(sides)) for _ in range(2))
        shapes.append((n0, n1))
        th, r, L = rng.uniform(0, np.pi), rng.uniform(1, 8), rng.uniform(0.02, 0.2)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        A = R @ np.diag([L, L / r])
        descs.append(pfb.make_desc(2, [n0, n1], [1.0, 1.0], covariance="gaussian",
                                   anisotropy="geometric", corr_length=list(A.ravel()),
                                   approximate=True, r2c_check_all=True, precision=precision))
        N0, N1, H, h = 2 * n0, 2 * n1, n0 + 1, n1 + 1
        falg.append(16 * N0 * h + 32 * H * h + 8 * H * N1 + 16 * H * n1 + 4 * n0 * n1)
    return descs, shapes, falg


def dctdst_1024(precision="fp32", cells=1024, seed=5):
    """1024^3 cells on [0,1]^3, Matern nu = 1.5 with full geometric
    anisotropy (a random rotation, lengths U[0.01, 0.05]), DCT matrix +
    DCT/DST field (the even-symmetric 2048^3 embedding), approximate."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    A = q @ np.diag(rng.uniform(0.01, 0.05, 3))
    return pfb.make_desc(3, [cells] * 3, [1.0] * 3, covariance="matern32",
                         anisotropy="geometric", corr_length=list(A.ravel()), matrix="dct",
                         field="dctdst", approximate=True, precision=precision)
