"""Shared helpers for the parity tests."""
import numpy as np


def rel(a, b):
    a = np.asarray(a).astype(np.complex128).reshape(-1)
    b = np.asarray(b).astype(np.complex128).reshape(-1)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def rand_traj(rng, d, M):
    return np.asfortranarray(rng.uniform(-0.5, 0.5, size=(d, M)).astype(np.float32))


def rand_img(rng, shape):
    return np.asfortranarray((rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64))
