"""TEST INFRASTRUCTURE -- FP64 restatement of the off-resonance (B0) model of
codeare's NFFT<cxfl> with "b0" and "timing" (NFFT.hpp:117-173; the NFFT3
mri_inh_3d plan of NFFTTraits.hpp:91-108,150-166,188-216), used only by tests
as the checker of the GPU path.

Model (b0 in cycles per unit of t, codeare's axis pairing):
    f_j = sum_r x(r) exp(-2 pi i k_j.r) exp(-2 pi i b0(r) t_j)

`ndft_b0` / `ndft_b0_adjoint` sum it directly.  `B0NFFT` restates the fast
algorithm -- the time-frequency term as a third NFFT dimension, after Knopp,
Kunis, Potts, "A note on the iterative MRI reconstruction from nonuniform
k-space data" (Int. J. Biomed. Imaging 2007) -- on the oracle's FP64 3D NFFT:
with ts the mid time, omega = B / (1/2 - m/N3) (B = max |b0|, the frequency
scale of NFFT.hpp:157), nu_r = N3 b0(r) / omega and tau_j = (t_j - ts) omega /
N3, the Kaiser-Bessel pair phi(u) (|u| < m) / phihat(tau) =
I0(m sqrt(b3^2 - (2 pi tau)^2)), b3 = pi (2 - 1/sigma3), sigma3 = 1.2 gives
    sum_l phi(l - nu) exp(-2 pi i l tau) ~= phihat(tau) exp(-2 pi i nu tau)
so f_j = NFFT3D[X](k_j, tau_j) / phihat(tau_j) with
X(r, l) = x(r) exp(-2 pi i b0(r) ts) phi(l - nu_r), l in [-N3/2, N3/2).
The size rule N3 >= 4 sigma3 T B + 2m (T half the time span, NFFT.hpp:150-155)
keeps |tau| <= 1/(2 sigma3) and every tap inside [-N3/2, N3/2); N3 is then
raised to an even length whose oversampled size has only the factors 2 and 3
(the GPU's FFT lengths), as the library does.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.special

from . import ora

SIGMA3 = 1.2


def _phase_matrix(k, N):
    """exp(-2 pi i k_j . r) for the column-major pixels of an (N0, N1) image."""
    k = np.asarray(k, dtype=np.float64)
    N0, N1 = N
    i0, i1 = np.meshgrid(np.arange(N0) - N0 // 2, np.arange(N1) - N1 // 2, indexing="ij")
    i0 = i0.reshape(-1, order="F")
    i1 = i1.reshape(-1, order="F")
    return np.exp(-2j * np.pi * (np.outer(k[1], i0) + np.outer(k[0], i1)))


def ndft_b0(k, t, b0, x):
    """Direct sum of the model (x, b0: (N0, N1) column-major)."""
    A = _phase_matrix(k, x.shape)
    B = np.exp(-2j * np.pi * np.outer(np.asarray(t, np.float64), np.asarray(b0, np.float64).reshape(-1, order="F")))
    return (A * B) @ np.asarray(x, np.complex128).reshape(-1, order="F")


def ndft_b0_adjoint(k, t, b0, f, N):
    A = _phase_matrix(k, N)
    B = np.exp(-2j * np.pi * np.outer(np.asarray(t, np.float64), np.asarray(b0, np.float64).reshape(-1, order="F")))
    return ((A * B).conj().T @ np.asarray(f, np.complex128)).reshape(N, order="F")


def _factors_23(n):
    for p in (2, 3):
        while n % p == 0:
            n //= p
    return n == 1


def third_axis(b0, t, m, sigma):
    """(N3, omega, ts) of the model: the size rule, then the FFT-length rule."""
    t = np.asarray(t, np.float64)
    B = float(np.max(np.abs(np.asarray(b0, np.float64))))
    ts = 0.5 * (t.min() + t.max())
    T = 0.5 * (t.max() - t.min())
    N3 = int(math.ceil(4.0 * SIGMA3 * T * B + 2.0 * m))
    N3 = max(N3 + (N3 & 1), 2 * m + 2)
    while True:
        n3 = int(math.ceil(sigma * N3))
        if n3 >= 2 * (2 * m + 2) and _factors_23(n3):
            break
        N3 += 2
    omega = B / (0.5 - m / N3) if B > 0 else 1.0
    return N3, omega, ts


def kb_phi(u, m, b):
    """Kaiser-Bessel window, truncated to |u| < m."""
    u = np.asarray(u, np.float64)
    s2 = m * m - u * u
    out = np.zeros_like(u)
    pos = s2 > 0
    s = np.sqrt(s2[pos])
    out[pos] = np.sinh(b * s) / (np.pi * s)
    return out


def kb_phihat(tau, m, b):
    a = 2.0 * np.pi * np.asarray(tau, np.float64)
    return scipy.special.i0(m * np.sqrt(np.maximum(b * b - a * a, 0.0)))


class B0NFFT:
    """The fast model on the oracle's FP64 3D NFFT (sigma, m as the plan's)."""

    def __init__(self, N, k, t, b0, sigma=2.0, m=3):
        self.N = tuple(int(v) for v in N)
        self.m = int(m)
        self.b0 = np.asarray(b0, np.float64)
        self.N3, self.omega, self.ts = third_axis(self.b0, t, m, sigma)
        self.b3 = math.pi * (2.0 - 1.0 / SIGMA3)
        t = np.asarray(t, np.float64)
        self.tau = (t - self.ts) * self.omega / self.N3
        self.nu = self.N3 * self.b0 / self.omega                       # (N0, N1)
        self.phase = np.exp(-2j * np.pi * self.b0 * self.ts)
        M = t.size
        k3 = np.zeros((3, M), dtype=np.float32)
        k3[0] = self.tau
        k3[1] = np.asarray(k)[0]
        k3[2] = np.asarray(k)[1]
        self.P = ora.NFFT(self.N + (self.N3,), M, sigma, m)
        self.P.kspace(np.asfortranarray(k3))
        l = np.arange(self.N3) - self.N3 // 2
        self.W = kb_phi(l[None, None, :] - self.nu[:, :, None], self.m, self.b3)   # (N0, N1, N3)
        self.phinv = 1.0 / kb_phihat(self.tau, self.m, self.b3)

    def trafo(self, x):
        X = (np.asarray(x, np.complex128) * self.phase)[:, :, None] * self.W
        return self.P.trafo(np.asfortranarray(X)) * self.phinv

    def adjoint(self, f):
        Y = self.P.adjoint(np.asarray(f, np.complex128) * self.phinv)
        return np.conj(self.phase) * np.sum(self.W * Y, axis=2)
