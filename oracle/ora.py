"""ctypes front end of the FP64 C oracle (``liboracle_cgs.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, as the checker / timed CPU port.  The product
(``codeare_amd``) never imports it.

Array conventions follow codeare (column-major ``Matrix``, Matrix.hpp:579), with
numpy arrays given in Fortran order of the codeare shape:
image (N0, N1[, N2]) complex, data (M, C) complex, maps (N0, N1[, N2], C),
trajectory (d, M) float32.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_cgs.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        P, D, F, I, I64 = ct.c_void_p, ct.c_double, ct.c_float, ct.c_int, ct.c_int64
        L.ora_nfft_create.restype = P
        L.ora_nfft_create.argtypes = [I, P, I64, D, I]
        L.ora_nfft_destroy.argtypes = [P]
        L.ora_nfft_n.restype = I
        L.ora_nfft_n.argtypes = [P, I]
        L.ora_nfft_set_trajectory.argtypes = [P, P]
        L.ora_nfft_trafo.argtypes = [P, P, P]
        L.ora_nfft_adjoint.argtypes = [P, P, P]
        L.ora_ndft_trafo.argtypes = [I, P, I64, P, P, P]
        L.ora_ndft_adjoint.argtypes = [I, P, I64, P, P, P]
        L.ora_bessel_i0.restype = D
        L.ora_bessel_i0.argtypes = [D]
        L.ora_kb_phi.restype = D
        L.ora_kb_phi.argtypes = [D, I, D]
        L.ora_kb_phihut.restype = D
        L.ora_kb_phihut.argtypes = [D, I, I, D]
        L.ora_fft1d.argtypes = [P, I, I]
        L.ora_damping_mask.argtypes = [I, P, P]
        L.ora_itrafo.restype = I
        L.ora_itrafo.argtypes = [P, P, P, P, I, D, P, P]
        L.ora_intensity_map.argtypes = [I64, I, P, P]
        L.ora_ncsense_E.argtypes = [P, I, P, P, P]
        L.ora_ncsense_EH_true.argtypes = [P, I, P, P, P, P]
        L.ora_ncsense_EH_ref.argtypes = [P, I, P, P, I, D, P, P]
        L.ora_cgls_ref.restype = I
        L.ora_cgls_ref.argtypes = [P, I, P, P, I, I, D, I, D, D, P, P, P]
        L.ora_cg_true.restype = I
        L.ora_cg_true.argtypes = [P, I, P, P, I, D, D, P, P, P, P]
        L.ora_cgls_ref_frames.restype = I
        L.ora_cgls_ref_frames.argtypes = [P, I, I, P, I64, P, P, I, I, D, I, D, D, P, P, P]
        L.ora_cg_true_frames.restype = I
        L.ora_cg_true_frames.argtypes = [P, I, I, P, I64, P, P, I, D, D, P, P, P]
        L.ora_time_marks.restype = I
        L.ora_time_marks.argtypes = [P, I]
        L.ora_set_threads.argtypes = [I]
        L.ora_get_threads.restype = I
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


def _c128(a):
    return np.asfortranarray(a, dtype=np.complex128)


def _c64(a):
    return np.asfortranarray(a, dtype=np.complex64)


def time_marks():
    """Seconds since the start of the last cg_true / cgls_ref call: [0] after
    E^H y, [i] after iteration i."""
    buf = np.zeros(4097, dtype=np.float64)
    n = lib().ora_time_marks(_ptr(buf), buf.size)
    return buf[:n]


def set_threads(n: int) -> None:
    lib().ora_set_threads(int(n))


def get_threads() -> int:
    return lib().ora_get_threads()


def bessel_i0(x: float) -> float:
    return lib().ora_bessel_i0(float(x))


def kb_phi(t: float, m: int, b: float) -> float:
    return lib().ora_kb_phi(float(t), int(m), float(b))


def kb_phihut(k: float, n: int, m: int, b: float) -> float:
    return lib().ora_kb_phihut(float(k), int(n), int(m), float(b))


def fft1d(x, sign: int):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().ora_fft1d(_ptr(a), a.size, int(sign))
    return a


def damping_mask(d: int, N) -> np.ndarray:
    N = list(N)
    arr = (ct.c_int * d)(*N)
    out = np.zeros([N[0]] * d, dtype=np.float64, order="F")
    lib().ora_damping_mask(d, ct.cast(arr, ct.c_void_p), _ptr(out))
    return out


def intensity_map(S) -> np.ndarray:
    S = _c64(S)
    C = S.shape[-1]
    nr = S.size // C
    out = np.zeros(S.shape[:-1], dtype=np.float32, order="F")
    lib().ora_intensity_map(nr, C, _ptr(S), _ptr(out))
    return out


def ndft_trafo(k, img):
    k = np.asfortranarray(k, dtype=np.float32)
    img = _c128(img)
    d, M = k.shape
    N = (ct.c_int * d)(*img.shape)
    f = np.zeros(M, dtype=np.complex128)
    lib().ora_ndft_trafo(d, ct.cast(N, ct.c_void_p), M, _ptr(k), _ptr(img), _ptr(f))
    return f


def ndft_adjoint(k, f, shape):
    k = np.asfortranarray(k, dtype=np.float32)
    f = _c128(f)
    d, M = k.shape
    N = (ct.c_int * d)(*shape)
    img = np.zeros(shape, dtype=np.complex128, order="F")
    lib().ora_ndft_adjoint(d, ct.cast(N, ct.c_void_p), M, _ptr(k), _ptr(f), _ptr(img))
    return img


class NFFT:
    """FP64 NFFT plan (NFFT<cxfl> semantics, NFFT.hpp:66-436)."""

    def __init__(self, N, M: int, sigma: float = 2.0, m: int = 3):
        self.N = tuple(int(v) for v in N)
        self.d = len(self.N)
        self.M = int(M)
        self.sigma = float(sigma)
        self.m = int(m)
        arr = (ct.c_int * self.d)(*self.N)
        self._p = lib().ora_nfft_create(self.d, ct.cast(arr, ct.c_void_p), self.M,
                                        self.sigma, self.m)
        if not self._p:
            raise ValueError("oracle: bad NFFT parameters")
        self.n = tuple(lib().ora_nfft_n(self._p, a) for a in range(self.d))
        self.k = None

    def __del__(self):
        p = getattr(self, "_p", None)
        if p and _lib is not None:
            _lib.ora_nfft_destroy(p)
            self._p = None

    def kspace(self, k):
        k = np.asfortranarray(k, dtype=np.float32)
        assert k.shape == (self.d, self.M)
        self.k = k
        lib().ora_nfft_set_trajectory(self._p, _ptr(k))

    def trafo(self, img):
        img = _c128(img)
        assert img.shape == self.N
        f = np.zeros(self.M, dtype=np.complex128)
        lib().ora_nfft_trafo(self._p, _ptr(img), _ptr(f))
        return f

    def adjoint(self, f):
        f = _c128(f).reshape(-1)
        assert f.size == self.M
        img = np.zeros(self.N, dtype=np.complex128, order="F")
        lib().ora_nfft_adjoint(self._p, _ptr(f), _ptr(img))
        return img

    def itrafo(self, y, w, maxit=3, eps=1e-3, what=None):
        y = _c128(y).reshape(-1)
        w = np.ascontiguousarray(w, dtype=np.float64)
        if what is None:
            what = damping_mask(self.d, self.N)
        what = np.asfortranarray(what, dtype=np.float64)
        out = np.zeros(self.N, dtype=np.complex128, order="F")
        hist = np.zeros(maxit + 1, dtype=np.float64)
        it = lib().ora_itrafo(self._p, _ptr(y), _ptr(w), _ptr(what), int(maxit),
                              float(eps), _ptr(out), _ptr(hist))
        return out, it, hist[: it + 1]

    # ---- NCSENSE level (complex64 I/O) ----------------------------------
    def E(self, S, x):
        S = _c64(S)
        x = _c64(x)
        C = S.shape[-1]
        y = np.zeros((self.M, C), dtype=np.complex64, order="F")
        lib().ora_ncsense_E(self._p, C, _ptr(S), _ptr(x), _ptr(y))
        return y

    def EH_true(self, S, w, y):
        S = _c64(S)
        y = _c64(y)
        w = np.ascontiguousarray(w, dtype=np.float32)
        C = S.shape[-1]
        x = np.zeros(self.N, dtype=np.complex64, order="F")
        lib().ora_ncsense_EH_true(self._p, C, _ptr(S), _ptr(w), _ptr(y), _ptr(x))
        return x

    def EH_ref(self, S, w, y, ftiter=3, fteps=1e-3):
        S = _c64(S)
        y = _c64(y)
        w = np.ascontiguousarray(w, dtype=np.float32)
        C = S.shape[-1]
        x = np.zeros(self.N, dtype=np.complex64, order="F")
        lib().ora_ncsense_EH_ref(self._p, C, _ptr(S), _ptr(w), int(ftiter),
                                 float(fteps), _ptr(y), _ptr(x))
        return x

    def cgls_ref(self, S, w, y, mode=1, ftiter=3, fteps=1e-3, maxit=10,
                 cgeps=1e-7, lam=0.0):
        S = _c64(S)
        y = _c64(y)
        w = np.ascontiguousarray(w, dtype=np.float32)
        C = S.shape[-1]
        x = np.zeros(self.N, dtype=np.complex64, order="F")
        hist = np.zeros(maxit + 1, dtype=np.float64)
        it = lib().ora_cgls_ref(self._p, C, _ptr(S), _ptr(w), int(mode), int(ftiter),
                                float(fteps), int(maxit), float(cgeps), float(lam),
                                _ptr(y), _ptr(x), _ptr(hist))
        return x, it, hist[: it + 1]

    def cg_true(self, S, w, y, maxit=50, tol=1e-5, lam=0.0, x0=None):
        S = _c64(S)
        y = _c64(y)
        w = np.ascontiguousarray(w, dtype=np.float32)
        C = S.shape[-1]
        x = np.zeros(self.N, dtype=np.complex64, order="F")
        hist = np.zeros(maxit + 1, dtype=np.float64)
        x0p = None
        if x0 is not None:
            x0 = _c64(x0)
            x0p = _ptr(x0)
        it = lib().ora_cg_true(self._p, C, _ptr(S), _ptr(w), int(maxit), float(tol),
                               float(lam), _ptr(y), x0p, _ptr(x), _ptr(hist))
        return x, it, hist[: it + 1]


# ---- frames / Cartesian slices solved together ------------------------------
def _frames_args(plans, S, w, y, maps_per_frame):
    F = len(plans)
    arr = (ct.c_void_p * F)(*[p._p for p in plans])
    moff = np.zeros(F + 1, dtype=np.int64)
    moff[1:] = np.cumsum([p.M for p in plans])
    S = _c64(S)
    C = S.shape[-1]
    Nt = int(np.prod(plans[0].N))
    if maps_per_frame:
        # codeare (N..., F, C) -> one (N..., C) block per frame
        S = _c64(np.moveaxis(S.reshape((Nt, F, C), order="F"), 1, 2))
        Sfs = Nt * C
    else:
        Sfs = 0
    w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
    y = _c64(y).reshape(-1, order="F")
    assert w.size == moff[-1] and y.size == moff[-1] * C
    return F, arr, moff, S, Sfs, C, w, y, Nt


def cgls_ref_frames(plans, S, w, y, mode=1, ftiter=3, fteps=1e-3, maxit=10, cgeps=1e-7, lam=0.0,
                    maps_per_frame=False):
    """codeare's CGLS on the whole (N^d, F) image of nmany frames / 3rd_dim_cart
    slices (one oracle NFFT plan per frame).  y holds the (M_f, C) block of every
    frame in turn; maps are (N..., C) shared or (N..., F, C) per frame."""
    F, arr, moff, S, Sfs, C, w, y, Nt = _frames_args(plans, S, w, y, maps_per_frame)
    x = np.zeros(plans[0].N + (F,), dtype=np.complex64, order="F")
    hist = np.zeros(maxit + 1, dtype=np.float64)
    it = lib().ora_cgls_ref_frames(ct.cast(arr, ct.c_void_p), F, C, _ptr(S), Sfs, _ptr(w), _ptr(moff),
                                   int(mode), int(ftiter), float(fteps), int(maxit), float(cgeps),
                                   float(lam), _ptr(y), _ptr(x), _ptr(hist))
    return x, it, hist[: it + 1]


def cg_true_frames(plans, S, w, y, maxit=50, tol=1e-5, lam=0.0, maps_per_frame=False):
    """The library's joint true-adjoint CG (FP64 scalars) on frames / slices."""
    F, arr, moff, S, Sfs, C, w, y, Nt = _frames_args(plans, S, w, y, maps_per_frame)
    x = np.zeros(plans[0].N + (F,), dtype=np.complex64, order="F")
    hist = np.zeros(maxit + 1, dtype=np.float64)
    it = lib().ora_cg_true_frames(ct.cast(arr, ct.c_void_p), F, C, _ptr(S), Sfs, _ptr(w), _ptr(moff),
                                  int(maxit), float(tol), float(lam), _ptr(y), _ptr(x), _ptr(hist))
    return x, it, hist[: it + 1]
