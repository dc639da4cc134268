"""Config-3 basket set-up shared by tests and bench.py: 16 GBM assets, S0 = 100,
sigma_a ~ U[0.15, 0.40] and a 3-factor correlation (off-diagonals ~0.3 +- 0.1),
both from numpy's PCG64 seeded with 0; 64 dates; equal weights; K = 100."""
import numpy as np


def c3_parameters(n=16, seed=0):
    rng = np.random.default_rng(seed)
    sig = rng.uniform(0.15, 0.40, n)
    B = rng.uniform(0.20, 0.43, (n, 3))
    C = B @ B.T
    np.fill_diagonal(C, 1.0)
    return sig, C


def cholesky_lower(C):
    return np.linalg.cholesky(C)
