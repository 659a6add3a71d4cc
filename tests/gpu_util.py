"""Helpers shared by the GPU tests (not collected: no test_ prefix)."""
import sys

import numpy as np


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_density(n, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    P = A @ A.T / n
    return P


def packed(F):
    return F[np.tril_indices(F.shape[0])].copy()
