"""TEST INFRASTRUCTURE -- FP64 direct-NUDFT restatement of the 2D CG-SENSE
normal equations, in torch (CPU or GPU), used only by tests to state the error
of the GPU path at full workload size, where the dense NDFT matrices of
oracle/ndft.py do not fit.

Conventions are codeare's (oracle/ndft.py): image axis a pairs with k(d-1-a, j)
and pixel i_a sits at frequency index i_a - N_a/2 (NFFT.hpp:267-268, 340-346);
images are (N0, N1) arrays indexed [i0, i1].

  E^H W y (i)   = sum_c conj(S_c(i)) sum_j w_j y_jc exp(+2 pi i phi_j(i))
  E^H W E x (i) = sum_c conj(S_c(i)) sum_i' K(i - i') S_c(i') x(i'),
  K(D)          = sum_j w_j exp(+2 pi i (D0 k(1,j) + D1 k(0,j)))
with phi_j(i) = (i0 - N0/2) k(1,j) + (i1 - N1/2) k(0,j).  K is summed directly
(no window, no grid) in FP64 and the convolution is applied exactly by a
zero-padded 2N FFT, which is exact for a linear convolution of support < 2N.
"""
from __future__ import annotations

import math

import torch

CHUNK = 8192


def _exp_outer(kv, idx):
    """exp(+2 pi i outer(kv, idx)) in complex128."""
    ph = torch.outer(kv, idx) * (2.0 * math.pi)
    return torch.polar(torch.ones_like(ph), ph)


def eh(k, w, y, S):
    """E^H W y.  k (2, M) float, w (M,), y (M, C) complex, S (N0, N1, C)."""
    dev = S.device
    k = torch.as_tensor(k, dtype=torch.float64, device=dev)
    w = torch.as_tensor(w, dtype=torch.float64, device=dev)
    y = torch.as_tensor(y, dtype=torch.complex128, device=dev)
    N0, N1, C = S.shape
    n0 = torch.arange(N0, dtype=torch.float64, device=dev) - N0 // 2
    n1 = torch.arange(N1, dtype=torch.float64, device=dev) - N1 // 2
    B = torch.zeros((N0, C * N1), dtype=torch.complex128, device=dev)
    M = k.shape[1]
    for s in range(0, M, CHUNK):
        e = min(M, s + CHUNK)
        u0 = _exp_outer(k[1, s:e], n0)                      # (m, N0)
        u1 = _exp_outer(k[0, s:e], n1)                      # (m, N1)
        z = (w[s:e, None] * y[s:e]).to(torch.complex128)    # (m, C)
        Z = (z[:, :, None] * u1[:, None, :]).reshape(e - s, C * N1)
        B += u0.transpose(0, 1) @ Z
    B = B.reshape(N0, C, N1).permute(0, 2, 1)               # (N0, N1, C)
    return torch.sum(torch.conj(S) * B, dim=-1)


def kernel_spectrum(k, w, N0, N1, device):
    """FFT of K(D) on the 2N x 2N grid, D stored at D mod 2N."""
    k = torch.as_tensor(k, dtype=torch.float64, device=device)
    w = torch.as_tensor(w, dtype=torch.float64, device=device)
    D0 = torch.arange(2 * N0, dtype=torch.float64, device=device) - N0
    D1 = torch.arange(2 * N1, dtype=torch.float64, device=device) - N1
    K = torch.zeros((2 * N0, 2 * N1), dtype=torch.complex128, device=device)
    M = k.shape[1]
    for s in range(0, M, CHUNK):
        e = min(M, s + CHUNK)
        v0 = _exp_outer(k[1, s:e], D0)
        v1 = _exp_outer(k[0, s:e], D1)
        K += v0.transpose(0, 1) @ (w[s:e, None].to(torch.complex128) * v1)
    K = torch.roll(K, shifts=(-N0, -N1), dims=(0, 1))
    return torch.fft.fft2(K)


def normal(x, S, Khat, lam=0.0):
    """(E^H W E + lam I) x, exactly (FP64)."""
    N0, N1, C = S.shape
    u = torch.zeros((C, 2 * N0, 2 * N1), dtype=torch.complex128, device=S.device)
    u[:, :N0, :N1] = (S * x[:, :, None]).permute(2, 0, 1)
    v = torch.fft.ifft2(torch.fft.fft2(u) * Khat[None])[:, :N0, :N1]
    return torch.sum(torch.conj(S).permute(2, 0, 1) * v, dim=0) + lam * x


def cg(S, Khat, b, lam=0.0, maxit=100, tol=0.0, x0=None, keep=None):
    """CG on (E^H W E + lam) x = b, FP64, norm-ratio stop ||r|| / ||b|| <= tol.
    Returns (x, iterations, residual history).  With keep (a dict whose keys
    are iteration counts) the iterate after each of those counts is stored
    under its key."""
    x = torch.zeros_like(b) if x0 is None else x0.clone().to(torch.complex128)
    r = b - normal(x, S, Khat, lam) if x0 is not None else b.clone()
    p = r.clone()
    bn = torch.vdot(b.flatten(), b.flatten()).real
    rn = torch.vdot(r.flatten(), r.flatten()).real
    hist = [math.sqrt(float(rn / bn))]
    it = 0
    while it < maxit and hist[-1] > tol:
        q = normal(p, S, Khat, lam)
        alpha = rn / torch.vdot(p.flatten(), q.flatten()).real
        x = x + alpha * p
        r = r - alpha * q
        rn_new = torch.vdot(r.flatten(), r.flatten()).real
        p = r + (rn_new / rn) * p
        rn = rn_new
        it += 1
        hist.append(math.sqrt(float(rn / bn)))
        if keep is not None and it in keep:
            keep[it] = x.clone()
    return x, it, hist
