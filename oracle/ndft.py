"""Dense FP64 twins of the path, in numpy, for tiny problems.

TEST INFRASTRUCTURE ONLY (see oracle/ora.py).  These pin the C oracle by
brute force: an explicit NDFT matrix with codeare's axis pairing
(NFFT.hpp:267-268 copies k(t,j) into NFFT3's x; NFFT.hpp:340-346 copies the
column-major image linearly into NFFT3's row-major f_hat, so image axis a pairs
with k(d-1-a, j)), and dense restatements of NFFT3's CGNR (NFFTTraits.hpp:122-138)
and CGLS (CGLS.hpp:30-59), in the spirit of the reference's dense-matrix twin
MCGLS (src/matrix/MCGLS.hpp:36-88).
"""
from __future__ import annotations

import numpy as np


def ndft_matrix(k: np.ndarray, N) -> np.ndarray:
    """A[j, i] = exp(-2 pi i sum_a (i_a - N_a/2) k(d-1-a, j)), i column-major."""
    k = np.asarray(k, dtype=np.float64)
    d, M = k.shape
    N = tuple(N)
    grids = np.meshgrid(*[np.arange(n) - n // 2 for n in N], indexing="ij")
    # column-major flattening of the image index
    idx = [g.reshape(-1, order="F") for g in grids]
    ph = np.zeros((M, idx[0].size))
    for a in range(d):
        ph += np.outer(k[d - 1 - a], idx[a])
    return np.exp(-2j * np.pi * ph)


def dense_E(k, N, S) -> np.ndarray:
    """E = [A diag(S_c)]_c stacked over coils; rows ordered (j fastest, c slowest)."""
    A = ndft_matrix(k, N)
    S = np.asarray(S, dtype=np.complex128)
    C = S.shape[-1]
    Sf = S.reshape(-1, C, order="F")
    return np.concatenate([A * Sf[:, c][None, :] for c in range(C)], axis=0)


def cgnr_dense(A, y, w, what, maxit, eps):
    """NFFT3 CGNR with PRECOMPUTE_WEIGHT|PRECOMPUTE_DAMP on an explicit matrix."""
    y = np.asarray(y, dtype=np.complex128)
    f = np.zeros(A.shape[1], dtype=np.complex128)
    r = y.copy()
    dot_r = np.sum(w * np.abs(r) ** 2)
    z = A.conj().T @ (w * r)
    dot_z = np.sum(what * np.abs(z) ** 2)
    p = z.copy()
    it = 0
    for _ in range(maxit):
        if dot_r < eps:
            break
        v = A @ (what * p)
        dot_v = np.sum(w * np.abs(v) ** 2)
        alpha = dot_z / dot_v
        f = f + alpha * what * p
        r = r - alpha * v
        dot_r = np.sum(w * np.abs(r) ** 2)
        z = A.conj().T @ (w * r)
        dz_old = dot_z
        dot_z = np.sum(what * np.abs(z) ** 2)
        beta = dot_z / dz_old
        p = beta * p + z
        it += 1
    return f, it


def cg_normal_dense(AH_A, b, lam, maxit, tol, x0=None):
    """CG on (AH_A + lam I) x = b with the norm-ratio stop ||r||/||b|| <= tol."""
    n = b.size
    Aop = AH_A + lam * np.eye(n)
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    r = b - Aop @ x
    p = r.copy()
    rn = np.vdot(r, r).real
    bn = np.vdot(b, b).real
    hist = [np.sqrt(rn / bn)]
    it = 0
    while hist[-1] > tol and it < maxit:
        q = Aop @ p
        alpha = rn / np.vdot(p, q).real
        x = x + alpha * p
        r = r - alpha * q
        rn_new = np.vdot(r, r).real
        p = r + (rn_new / rn) * p
        rn = rn_new
        it += 1
        hist.append(np.sqrt(rn / bn))
    return x, it, np.array(hist)
