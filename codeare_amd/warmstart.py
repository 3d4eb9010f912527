"""Learned warm start for the CG-SENSE solve (no counterpart in codeare, whose
CGLS starts from zero: src/optimisation/CGLS.hpp:30-59).

A 2D U-Net (include/cgs_unet.h, BF16 MFMA kernels in csrc/unet.hip) maps the
scaled adjoint image x_a = alpha0 * E^H W y, alpha0 = <b,b> / <b, A b> (the
first steepest-descent step of the normal equations A x = b, A = E^H W E +
lambda I), to the converged CG image.  Inputs and targets are divided by the
RMS of x_a so the network sees unit-scale images; the head adds its output to
x_a (residual form), so the untrained network returns x_a itself.

The solve from the network's output uses the same stopping test as the solve
from zero (||r|| / ||E^H W y|| <= tol); zero-start CG stays the fallback.
"""
from __future__ import annotations

import ctypes as ct
import time

import numpy as np

from . import cgsense as _c
from .cgsense import _check, _ptr

UNET_EXPORTED = [
    "cgs_unet_create", "cgs_unet_destroy", "cgs_unet_num_params", "cgs_unet_get_params",
    "cgs_unet_set_params", "cgs_unet_forward", "cgs_unet_gradients", "cgs_unet_get_grads",
    "cgs_unet_train_step", "cgs_unet_reset_optimizer", "cgs_unet_profile",
]

_bound = False


def _lib():
    global _bound
    L = _c.load_library()
    if not _bound:
        P, I, D = ct.c_void_p, ct.c_int, ct.c_double
        L.cgs_unet_create.argtypes = [ct.POINTER(P), I, I, I, I, I, ct.c_ulonglong]
        L.cgs_unet_destroy.argtypes = [P]
        L.cgs_unet_num_params.argtypes = [P, ct.POINTER(ct.c_longlong)]
        L.cgs_unet_get_params.argtypes = [P, P]
        L.cgs_unet_set_params.argtypes = [P, P]
        L.cgs_unet_forward.argtypes = [P, P, P, I, I]
        L.cgs_unet_gradients.argtypes = [P, P, P, I, I, ct.POINTER(D)]
        L.cgs_unet_get_grads.argtypes = [P, P]
        L.cgs_unet_train_step.argtypes = [P, P, P, I, D, I, ct.POINTER(D)]
        L.cgs_unet_reset_optimizer.argtypes = [P]
        L.cgs_unet_profile.argtypes = [P, I, P]
        for n in UNET_EXPORTED:
            getattr(L, n).restype = I
        _bound = True
    return L


class UNet:
    """The warm-start network.  Images are complex64 arrays (B, H, W) or
    codeare-layout (N0, N1) Fortran images (see `images_to_batch`)."""

    def __init__(self, H: int, W: int, levels: int = 4, base_channels: int = 32,
                 max_batch: int = 16, seed: int = 1):
        L = _lib()
        self._h = ct.c_void_p()
        _check(L.cgs_unet_create(ct.byref(self._h), H, W, levels, base_channels, max_batch, seed))
        self.H, self.W, self.levels, self.max_batch = H, W, levels, max_batch
        n = ct.c_longlong()
        _check(L.cgs_unet_num_params(self._h, ct.byref(n)))
        self.nparams = int(n.value)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib().cgs_unet_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ---------------------------------------------------------------
    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(_lib().cgs_unet_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, theta) -> None:
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {theta.size}")
        _check(_lib().cgs_unet_set_params(self._h, _ptr(theta)))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(_lib().cgs_unet_get_grads(self._h, _ptr(out)))
        return out

    def save(self, path: str) -> None:
        np.savez(path, theta=self.params(), H=self.H, W=self.W, levels=self.levels)

    def load(self, path: str) -> None:
        z = np.load(path, allow_pickle=False)
        if int(z["H"]) != self.H or int(z["W"]) != self.W or int(z["levels"]) != self.levels:
            raise ValueError("saved network has another shape")
        self.set_params(z["theta"])

    # compute ------------------------------------------------------------------
    def _batch(self, x):
        if hasattr(x, "data_ptr"):
            return x, int(x.shape[0]), 1
        x = np.ascontiguousarray(x, dtype=np.complex64)
        if x.ndim == 2:
            x = x[None]
        return x, int(x.shape[0]), 0

    def forward(self, x, out=None):
        x, B, dev = self._batch(x)
        if out is None:
            out = np.empty((B, self.H, self.W), dtype=np.complex64)
        _check(_lib().cgs_unet_forward(self._h, _ptr(x), _ptr(out), B, dev))
        return out

    def gradients(self, x, t) -> float:
        x, B, dev = self._batch(x)
        t, _, _ = self._batch(t)
        loss = ct.c_double()
        _check(_lib().cgs_unet_gradients(self._h, _ptr(x), _ptr(t), B, dev, ct.byref(loss)))
        return float(loss.value)

    def train_step(self, x, t, lr: float = 1e-3) -> float:
        x, B, dev = self._batch(x)
        t, _, _ = self._batch(t)
        loss = ct.c_double()
        _check(_lib().cgs_unet_train_step(self._h, _ptr(x), _ptr(t), B, float(lr), dev, ct.byref(loss)))
        return float(loss.value)

    def reset_optimizer(self) -> None:
        _check(_lib().cgs_unet_reset_optimizer(self._h))

    def profile(self, enable: int = -1) -> dict:
        s = np.zeros(8)
        _check(_lib().cgs_unet_profile(self._h, int(enable), _ptr(s)))
        return {"fwd_ms": s[0], "fwd_flop": s[1], "dgrad_ms": s[2], "dgrad_flop": s[3],
                "wgrad_ms": s[4], "wgrad_flop": s[5], "other_ms": s[6]}


def images_to_batch(imgs) -> np.ndarray:
    """codeare (N0, N1) Fortran images -> (B, H=N1, W=N0) C-order batch (same bytes)."""
    return np.stack([np.asarray(im, dtype=np.complex64).T for im in imgs])


def batch_to_image(b) -> np.ndarray:
    return np.asfortranarray(np.asarray(b).T)


# ---------------------------------------------------------------------------
# solver glue
def network_input(plan, y, lam: float):
    """(x_a, scale, b): x_a = alpha0 * b with b = E^H W y, alpha0 = <b,b>/<b,Ab>,
    and scale = RMS(x_a) (network input is x_a / scale)."""
    b = plan.EH(y)
    Ab = plan.normal(b, lam=lam, toeplitz=True)
    bb = float(np.vdot(b, b).real)
    bAb = float(np.vdot(b, Ab).real)
    xa = (bb / bAb) * b if bAb > 0 else b
    scale = float(np.sqrt(np.mean(np.abs(xa.astype(np.complex128)) ** 2))) or 1.0
    return xa, scale, b


def warm_start(net: UNet, plan, y, lam: float):
    """x0 from the network, plus the time it took (s)."""
    t0 = time.perf_counter()
    xa, scale, _ = network_input(plan, y, lam)
    out = net.forward(images_to_batch([xa / scale]))
    x0 = batch_to_image(out[0]) * np.float32(scale)
    return x0.astype(np.complex64), time.perf_counter() - t0
