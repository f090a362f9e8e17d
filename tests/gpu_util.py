"""Helpers for the GPU parity tests."""
import numpy as np


def relerr(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    n = np.linalg.norm(ref.ravel())
    return np.linalg.norm((x - ref).ravel()) / (n if n > 0 else 1.0)


def perturb(P, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    P.udg = np.asfortranarray(P.udg * (1 + amp * rng.standard_normal(P.udg.shape)))
    P.uh = np.asfortranarray(P.uh * (1 + amp * rng.standard_normal(P.uh.shape)))
    return P
