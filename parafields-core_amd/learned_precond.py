"""Learned spectral preconditioner for multiplyInverse (no reference
counterpart).

A small MLP (6 inputs -> 3 x 64 exact-GELU hidden -> 64 outputs, FP32) maps
covariance parameters to per-radial-frequency-bin log-multipliers L_b. The
conjugate-gradient solve of (C/N_ext) x = b (Matrix::multiplyInverse,
include/parafields/matrix.hh:189-206) is then preconditioned by
M^{-1} = R F^{-1} diag(exp(L_bin(k))) F E on the embedding
(PF_CG_LEARNED_PCG). Plain CG remains the fallback, and the FP64 true
residual judges convergence.

Training data come from the method itself. For each parameter set, the
library computes the FP64 embedding eigenvalues, and pf_spectral_targets
reduces them to L*_b = -log(mean over bin b of max(lambda_k, nugget/N_ext)).
Training (forward, backward, Adam) runs in the library's HIP kernels.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

import pf_binding as pfb

NBINS = 64
DIMS = [6, 64, 64, 64, NBINS]
SMOOTHNESS = {"exponential": 0.5, "matern32": 1.5, "matern52": 2.5, "gaussian": 5.0,
              "custom-iso": 0.5}


@dataclass
class CovCase:
    """One parameter set: the inputs of the network."""
    variance: float
    corr_length: Sequence[float]
    covariance: str = "exponential"
    cells: int = 128
    dim: int = 3
    factor: int = 2
    nugget: float = 1e-3

    def desc(self, precision: str = "fp64", cg_iterations: int = 100000):
        cl = list(self.corr_length)
        aniso = "none" if len(cl) == 1 else "axiparallel"
        return pfb.make_desc(self.dim, [self.cells] * self.dim, [1.0] * self.dim,
                             covariance=self.covariance, variance=self.variance,
                             corr_length=cl, anisotropy=aniso, factor=self.factor,
                             nugget=self.nugget, approximate=True, precision=precision,
                             cg_iterations=cg_iterations)

    def features(self) -> np.ndarray:
        cl = np.asarray(self.corr_length, dtype=np.float64)
        return np.array([
            math.log(self.variance),
            float(np.log(cl).mean()),
            SMOOTHNESS.get(self.covariance, 1.0),
            float(np.log(cl.max() / cl.min())),
            math.log(self.cells) / 5.0,
            self.factor / 2.0,
        ], dtype=np.float32)


def sample_cases(n: int, seed: int, cells: int = 128, lmin: float = 0.05, lmax: float = 0.3,
                 vmin: float = 0.5, vmax: float = 2.0) -> List[CovCase]:
    """Exponential covariance, l ~ U[lmin, lmax], variance ~ U[vmin, vmax], nugget 1e-3."""
    rng = np.random.default_rng(seed)
    return [CovCase(variance=float(rng.uniform(vmin, vmax)),
                    corr_length=[float(rng.uniform(lmin, lmax))], cells=cells)
            for _ in range(n)]


def method_targets(case: CovCase, nbins: int = NBINS) -> np.ndarray:
    """L*_b from the method's own FP64 eigenvalues (one plan per case)."""
    d = case.desc()
    plan = pfb.Plan(d)
    try:
        floor = case.nugget / float(plan.sizes.extended_size)
        return plan.spectral_targets(nbins, floor)
    finally:
        plan.close()


@dataclass
class Model:
    mlp: pfb.Mlp
    mu: np.ndarray
    sigma: np.ndarray

    def predict(self, cases: Sequence[CovCase]) -> np.ndarray:
        x = np.stack([c.features() for c in cases])
        y = self.mlp.forward(x).astype(np.float64)
        return y * self.sigma + self.mu

    def save(self, path: str):
        np.savez(path, params=self.mlp.params(), mu=self.mu, sigma=self.sigma,
                 dims=np.asarray(self.mlp.dims))

    @staticmethod
    def load(path: str) -> "Model":
        z = np.load(path)
        m = pfb.Mlp([int(v) for v in z["dims"]])
        m.set_params(z["params"])
        return Model(m, z["mu"], z["sigma"])


def build_training_set(cases: Sequence[CovCase], nbins: int = NBINS):
    x = np.stack([c.features() for c in cases])
    y = np.stack([method_targets(c, nbins) for c in cases])
    return x, y


def train(x: np.ndarray, y: np.ndarray, steps: int = 3000, lr: float = 2e-3, seed: int = 0,
          chunks: int = 6) -> tuple:
    """Standardise the targets per bin and fit the MLP with full-batch Adam on
    the device; the learning rate halves at each chunk after the first."""
    mu = y.mean(axis=0)
    sigma = y.std(axis=0) + 1e-6
    yn = ((y - mu) / sigma).astype(np.float32)
    mlp = pfb.Mlp([x.shape[1]] + DIMS[1:-1] + [y.shape[1]], seed=seed)
    losses = []
    per = max(1, steps // chunks)
    for c in range(chunks):
        losses.append(mlp.train(x, yn, per, lr * (0.5 ** max(0, c - 1))))
    return Model(mlp, mu, sigma), losses


def solve(case: CovCase, b: np.ndarray, mode: str, model: Model | None = None,
          rel_tol: float = 1e-8, max_iterations: int = 20000):
    """multiplyInverse of `b` by (P)CG on the device; returns (x, stats).
    Modes: "plain", "embed_pcg" (the embedding inverse, the reference's
    preconditioner), "learned_pcg" (the network's multipliers) and
    "spectral_pcg" (the same radial preconditioner with the exact targets of
    this case, i.e. what a perfect network would predict)."""
    d = case.desc(cg_iterations=max_iterations)
    plan = pfb.Plan(d)
    try:
        lm = None
        if mode == "learned_pcg":
            lm = model.predict([case])[0].astype(np.float32)
        elif mode == "spectral_pcg":
            floor = case.nugget / float(plan.sizes.extended_size)
            lm = plan.spectral_targets(NBINS, floor).astype(np.float32)
            mode = "learned_pcg"
        x, st = plan.cg(b, mode=mode, rel_tol=rel_tol, max_iterations=max_iterations,
                        bin_log_multipliers=lm)
        return x, st
    finally:
        plan.close()
