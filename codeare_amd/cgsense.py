"""Python host mirror of codeare's NFFT<cxfl> / NCSENSE<cxfl> operators over the
C ABI of libcgsense_hip.so (include/cgsense.h).

The classes keep the reference's interface for this path:
  NFFT(params)     -> KSpace, Weights, Trafo (*), Adjoint (->*, the iterative
                      NFFT3 inverse of NFFT.hpp:370-436), plus adjoint_true
  NCSENSE(params)  -> KSpace, Weights, Sensitivities, Trafo (*), operator/ (/),
                      Adjoint (->*, the CGLS solve of NCSENSE.hpp:408-414)
with codeare's Params keys and defaults (NCSENSE.hpp:69-184, NFFT.hpp:66-191).
Arrays are numpy, column-major (Fortran order) in codeare's shapes: images
(N0, N1[, N2]), data (M, C), maps (N0, N1[, N2], C), trajectory (d, M).

The HIP library is required: if it is missing, loading raises; there is no CPU
fallback in this package.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcgsense_hip.so")

CGS_OK = 0
CGS_ERR_ARG = 1
CGS_ERR_STATE = 2
CGS_ERR_HIP = 3
CGS_ERR_ALLOC = 4
CGS_ERR_UNSUPPORTED = 5
CGS_ERR_NAN = 6
CGS_DEVICE_IO = 0x1
CGS_PROFILE = 0x2
CGS_DETERMINISTIC = 0x4
CGS_AUTOTUNE = 0x8
CGS_MAPS_PER_FRAME = 0x10
CGS_EH_TRUE = 0
CGS_EH_REF = 1
CGS_TOL_NORM_RATIO = 0
CGS_TOL_SQUARED_RATIO = 1
CGS_NORMAL_NUFFT = 0
CGS_NORMAL_TOEPLITZ = 1
KERNELS = ["fft_fwd", "interp", "grid", "fft_adj", "coil", "cg", "other", "toep_pad", "toep_conv", "toep_crop"]

EXPORTED = [
    "cgs_plan_create", "cgs_plan_destroy", "cgs_set_stream", "cgs_sync",
    "cgs_set_trajectory", "cgs_set_weights", "cgs_set_sensitivities", "cgs_set_sensitivities_device",
    "cgs_nfft_forward", "cgs_nfft_adjoint", "cgs_nfft_inverse_ref",
    "cgs_E", "cgs_EH", "cgs_normal", "cgs_solve", "cgs_solve_noalloc",
    "cgs_kernel_stats", "cgs_plan_info", "cgs_status_string", "cgs_last_error",
    "cgs_residual_history", "cgs_autotune", "cgs_profile_mask", "cgs_set_b0", "cgs_b0_info",
    "cgs_gridding_kind", "cgs_time_toeplitz_conv",
]


class CgsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cgsense error {code}: {msg}")
        self.code = code


class SolveOpts(ct.Structure):
    _fields_ = [("maxit", ct.c_int), ("tol", ct.c_double), ("tol_kind", ct.c_int),
                ("lambda_", ct.c_double), ("eh_mode", ct.c_int), ("ftiter", ct.c_int),
                ("fteps", ct.c_double), ("normal_op", ct.c_int), ("graph_iters", ct.c_int),
                ("joint", ct.c_int)]


class TuneReport(ct.Structure):
    _fields_ = [("visits_per_item", ct.c_int), ("grid_coils", ct.c_int), ("interp_coils", ct.c_int),
                ("from_cache", ct.c_int), ("grid_ms", ct.c_double), ("interp_ms", ct.c_double),
                ("default_grid_ms", ct.c_double), ("default_interp_ms", ct.c_double),
                ("grid_mfma", ct.c_int), ("grid_sweep", ct.c_int), ("grid_bin", ct.c_int),
                ("interp_lanes", ct.c_int)]


class SolveReport(ct.Structure):
    _fields_ = [("iters", ct.c_int * 1024), ("final_rel", ct.c_double * 1024),
                ("nframes", ct.c_int), ("max_iters", ct.c_int), ("status", ct.c_int),
                ("done_ms", ct.c_double * 1024), ("total_ms", ct.c_double)]


_lib = None


def load_library(path: str | None = None):
    """Load libcgsense_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with codeare_amd/build.py (hipcc, gfx950)")
    L = ct.CDLL(path)
    P, I, D, I64 = ct.c_void_p, ct.c_int, ct.c_double, ct.c_int64
    L.cgs_plan_create.argtypes = [ct.POINTER(P), I, P, I64, D, I, I, I, P, ct.c_uint]
    L.cgs_plan_destroy.argtypes = [P]
    L.cgs_set_stream.argtypes = [P, P]
    L.cgs_sync.argtypes = [P]
    L.cgs_set_trajectory.argtypes = [P, P]
    L.cgs_set_weights.argtypes = [P, P]
    L.cgs_set_sensitivities.argtypes = [P, P]
    L.cgs_set_sensitivities_device.argtypes = [P, P]
    L.cgs_gridding_kind.argtypes = [P]
    L.cgs_time_toeplitz_conv.argtypes = [P, I, ct.POINTER(D)]
    L.cgs_nfft_forward.argtypes = [P, I, P, P, I]
    L.cgs_nfft_adjoint.argtypes = [P, I, P, P, I]
    L.cgs_nfft_inverse_ref.argtypes = [P, I, P, P, I, I, D]
    L.cgs_E.argtypes = [P, P, P]
    L.cgs_EH.argtypes = [P, P, P, I, I, D]
    L.cgs_normal.argtypes = [P, P, P, D, I]
    L.cgs_solve.argtypes = [P, P, P, P, ct.POINTER(SolveOpts), ct.POINTER(SolveReport)]
    L.cgs_solve_noalloc.argtypes = [P, P, P, P, ct.POINTER(SolveOpts), ct.POINTER(SolveReport)]
    L.cgs_kernel_stats.argtypes = [P, I, ct.POINTER(I64), ct.POINTER(D), ct.POINTER(D), I]
    L.cgs_plan_info.argtypes = [P, ct.POINTER(I), ct.POINTER(I), ct.POINTER(I64), ct.POINTER(I)]
    L.cgs_residual_history.argtypes = [P, I, P, I]
    L.cgs_autotune.argtypes = [P, I, I, ct.POINTER(TuneReport)]
    L.cgs_profile_mask.argtypes = [P, ct.c_uint]
    L.cgs_set_b0.argtypes = [P, P, P]
    L.cgs_b0_info.argtypes = [P, ct.POINTER(I), ct.POINTER(D), ct.POINTER(D)]
    L.cgs_status_string.restype = ct.c_char_p
    L.cgs_status_string.argtypes = [I]
    L.cgs_last_error.restype = ct.c_char_p
    for name in EXPORTED:
        if name not in ("cgs_status_string", "cgs_last_error"):
            getattr(L, name).restype = I
    _lib = L
    return L


def _check(rc: int):
    if rc != CGS_OK:
        L = load_library()
        raise CgsError(rc, f"{L.cgs_status_string(rc).decode()}: {L.cgs_last_error().decode()}")


def _ptr(a):
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ct.c_void_p)
    # torch tensor or raw integer pointer (device memory)
    if hasattr(a, "data_ptr"):
        return ct.c_void_p(a.data_ptr())
    return ct.c_void_p(int(a))


def _c64(a):
    return np.asfortranarray(a, dtype=np.complex64)


class Plan:
    """One trajectory (optionally several frames), weights, maps and buffers."""

    def __init__(self, N, M_total: int, sigma: float = 2.0, m: int = 3, ncoils: int = 1,
                 frame_offsets=None, device_io: bool = False, profile: bool = False,
                 deterministic: bool = False, autotune: bool = False, maps_per_frame: bool = False):
        L = load_library()
        self.N = tuple(int(v) for v in N)
        self.d = len(self.N)
        self.M = int(M_total)
        self.C = int(ncoils)
        self.sigma = float(sigma)
        self.m = int(m)
        if frame_offsets is None:
            frame_offsets = np.array([0, self.M], dtype=np.int64)
        self.foff = np.ascontiguousarray(frame_offsets, dtype=np.int64)
        self.F = len(self.foff) - 1
        self.device_io = device_io
        self.maps_per_frame = bool(maps_per_frame)
        flags = ((CGS_DEVICE_IO if device_io else 0) | (CGS_PROFILE if profile else 0)
                 | (CGS_DETERMINISTIC if deterministic else 0) | (CGS_AUTOTUNE if autotune else 0)
                 | (CGS_MAPS_PER_FRAME if maps_per_frame else 0))
        Narr = (ct.c_int * self.d)(*self.N)
        h = ct.c_void_p()
        _check(L.cgs_plan_create(ct.byref(h), self.d, ct.cast(Narr, ct.c_void_p), self.M,
                                 self.sigma, self.m, self.C, self.F,
                                 self.foff.ctypes.data_as(ct.c_void_p) if self.F > 1 else None,
                                 flags))
        self._h = h
        self.Ntot = int(np.prod(self.N))

    def close(self):
        if getattr(self, "_h", None):
            load_library().cgs_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- setup -------------------------------------------------------------
    def set_stream(self, stream_ptr: int):
        _check(load_library().cgs_set_stream(self._h, ct.c_void_p(stream_ptr)))

    def sync(self):
        _check(load_library().cgs_sync(self._h))

    def kspace(self, k):
        k = np.asfortranarray(k, dtype=np.float32)
        if k.shape != (self.d, self.M):
            raise ValueError(f"trajectory must be ({self.d}, {self.M})")
        self._k = k
        _check(load_library().cgs_set_trajectory(self._h, _ptr(k)))

    def weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        if w.size != self.M:
            raise ValueError("weights must have M_total entries")
        _check(load_library().cgs_set_weights(self._h, _ptr(w)))

    def sensitivities(self, S):
        S = _c64(S)
        nm = self.F if self.maps_per_frame else 1
        if S.size != self.Ntot * self.C * nm:
            raise ValueError("maps must be (N..., C)" if nm == 1 else "maps must be (N..., F, C)")
        self._S = S
        _check(load_library().cgs_set_sensitivities(self._h, _ptr(S)))

    def sensitivities_device(self, S):
        """Maps already on the device (a torch tensor, the layout of
        `sensitivities`), loaded on the plan's stream; keep S alive until then."""
        nm = self.F if self.maps_per_frame else 1
        if S.numel() != self.Ntot * self.C * nm or not S.is_cuda:
            raise ValueError("maps must be a device tensor of (N..., C) entries")
        _check(load_library().cgs_set_sensitivities_device(self._h, ct.c_void_p(S.data_ptr())))

    def b0(self, b0, t):
        """Off-resonance model of the NFFT operators (codeare's NFFT "b0" and
        "timing", NFFT.hpp:117-173): b0 (N0, N1) in cycles per unit of t,
        t (M,) sample times; None removes it."""
        L = load_library()
        if b0 is None:
            _check(L.cgs_set_b0(self._h, None, None))
            return
        b0 = np.asfortranarray(b0, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        if b0.size != self.Ntot or t.size != self.M:
            raise ValueError("b0 must have N0*N1 entries and timing M_total")
        _check(L.cgs_set_b0(self._h, _ptr(b0), _ptr(t)))

    def b0_info(self):
        n3 = ct.c_int()
        om = ct.c_double()
        ts = ct.c_double()
        _check(load_library().cgs_b0_info(self._h, ct.byref(n3), ct.byref(om), ct.byref(ts)))
        return {"N3": n3.value, "omega": om.value, "ts": ts.value}

    def info(self):
        n = (ct.c_int * 3)()
        t = ct.c_int()
        v = ct.c_int64()
        gi = ct.c_int()
        _check(load_library().cgs_plan_info(self._h, n, ct.byref(t), ct.byref(v), ct.byref(gi)))
        return {"n": tuple(n)[: self.d], "tile": t.value, "visits": v.value, "grid_items": gi.value}

    # ---- operators (numpy host I/O unless device_io) ------------------------
    def _frame_M(self, frame):
        return int(self.foff[frame + 1] - self.foff[frame])

    def nfft_forward(self, img, frame: int = 0, out=None):
        L = load_library()
        if self.device_io:
            batch = img.numel() // self.Ntot
            _check(L.cgs_nfft_forward(self._h, frame, _ptr(img), _ptr(out), batch))
            return out
        img = _c64(img)
        batch = img.size // self.Ntot
        f = np.zeros((self._frame_M(frame), batch), dtype=np.complex64, order="F")
        _check(L.cgs_nfft_forward(self._h, frame, _ptr(img), _ptr(f), batch))
        return f[:, 0] if img.ndim == self.d else f

    def nfft_adjoint(self, f, frame: int = 0, out=None):
        L = load_library()
        f = _c64(f)
        batch = f.size // self._frame_M(frame)
        img = np.zeros(self.N + ((batch,) if batch > 1 else ()), dtype=np.complex64, order="F")
        _check(L.cgs_nfft_adjoint(self._h, frame, _ptr(f), _ptr(img), batch))
        return img

    def nfft_inverse_ref(self, f, frame: int = 0, maxit: int = 3, eps: float = 7e-4):
        L = load_library()
        f = _c64(f)
        batch = f.size // self._frame_M(frame)
        img = np.zeros(self.N + ((batch,) if batch > 1 else ()), dtype=np.complex64, order="F")
        _check(L.cgs_nfft_inverse_ref(self._h, frame, _ptr(f), _ptr(img), batch, int(maxit), float(eps)))
        return img

    def _img_shape(self):
        return self.N + ((self.F,) if self.F > 1 else ())

    def E(self, x, out=None):
        L = load_library()
        if self.device_io:
            _check(L.cgs_E(self._h, _ptr(x), _ptr(out)))
            return out
        x = _c64(x)
        y = np.zeros(self.M * self.C, dtype=np.complex64)
        _check(L.cgs_E(self._h, _ptr(x), _ptr(y)))
        if self.F == 1:
            return y.reshape((self.M, self.C), order="F")
        return y

    def EH(self, y, mode: int = CGS_EH_TRUE, ftiter: int = 3, fteps: float = 1e-3, out=None):
        L = load_library()
        if self.device_io:
            _check(L.cgs_EH(self._h, _ptr(y), _ptr(out), int(mode), int(ftiter), float(fteps)))
            return out
        y = _c64(y)
        x = np.zeros(self._img_shape(), dtype=np.complex64, order="F")
        _check(L.cgs_EH(self._h, _ptr(y), _ptr(x), int(mode), int(ftiter), float(fteps)))
        return x

    def normal(self, x, lam: float = 0.0, out=None, toeplitz: bool = False):
        L = load_library()
        op = CGS_NORMAL_TOEPLITZ if toeplitz else CGS_NORMAL_NUFFT
        if self.device_io:
            _check(L.cgs_normal(self._h, _ptr(x), _ptr(out), float(lam), op))
            return out
        x = _c64(x)
        q = np.zeros(self._img_shape(), dtype=np.complex64, order="F")
        _check(L.cgs_normal(self._h, _ptr(x), _ptr(q), float(lam), op))
        return q

    def solve(self, y, x0=None, maxit: int = 10, tol: float = 1e-5, tol_kind: int = CGS_TOL_NORM_RATIO,
              lam: float = 0.0, eh_mode: int = CGS_EH_TRUE, ftiter: int = 3, fteps: float = 1e-3,
              out=None, check_nan: bool = True, toeplitz: bool = False, graph_iters: int = 0,
              joint: bool = False, noalloc: bool = False):
        """CG solve (cgs_solve); with noalloc the allocation-free variant
        cgs_solve_noalloc (device-I/O plans, graph_iters 0)."""
        L = load_library()
        fn = L.cgs_solve_noalloc if noalloc else L.cgs_solve
        o = SolveOpts(int(maxit), float(tol), int(tol_kind), float(lam), int(eh_mode), int(ftiter),
                      float(fteps), CGS_NORMAL_TOEPLITZ if toeplitz else CGS_NORMAL_NUFFT,
                      int(graph_iters), int(bool(joint)))
        rep = SolveReport()
        if self.device_io:
            rc = fn(self._h, _ptr(y), _ptr(x0) if x0 is not None else None, _ptr(out),
                    ct.byref(o), ct.byref(rep))
            x = out
        else:
            y = _c64(y)
            x = np.zeros(self._img_shape(), dtype=np.complex64, order="F")
            x0p = None
            if x0 is not None:
                x0 = _c64(x0)
                x0p = _ptr(x0)
            rc = fn(self._h, _ptr(y), x0p, _ptr(x), ct.byref(o), ct.byref(rep))
        if rc != CGS_ERR_NAN or check_nan:
            _check(rc)
        F = min(self.F, 1024)
        report = {"iters": list(rep.iters[:F]), "final_rel": list(rep.final_rel[:F]),
                  "max_iters": rep.max_iters, "status": rep.status,
                  "done_ms": list(rep.done_ms[:F]), "total_ms": rep.total_ms}
        return x, report

    def time_toeplitz_conv(self, reps: int = 200) -> float:
        """Average µs of the Toeplitz convolution pass, `reps` back-to-back launches."""
        us = ct.c_double()
        _check(load_library().cgs_time_toeplitz_conv(self._h, int(reps), ct.byref(us)))
        return us.value

    def gridding_kind(self) -> str:
        """The scatter the plan runs now: "tiles", "sweep" or "binned"."""
        k = load_library().cgs_gridding_kind(self._h)
        if k < 0:
            _check(-k)
        return {0: "tiles", 1: "sweep", 2: "binned"}[k]

    def autotune(self, reps: int = 5, use_cache: bool = False) -> dict:
        """Time the gridding/gather launch shapes on this trajectory, keep the fastest."""
        r = TuneReport()
        _check(load_library().cgs_autotune(self._h, int(reps), int(use_cache), ct.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneReport._fields_}

    def residual_history(self, frame: int = 0, cap: int = 100000):
        buf = np.zeros(cap, dtype=np.float64)
        n = load_library().cgs_residual_history(self._h, int(frame), _ptr(buf), int(cap))
        if n < 0:
            _check(-n)
        return buf[:n]

    def profile_mask(self, kernels=None):
        """Time only these kernel classes (names from KERNELS); None = all."""
        mask = 0xFFFFFFFF if kernels is None else sum(1 << KERNELS.index(k) for k in kernels)
        _check(load_library().cgs_profile_mask(self._h, ct.c_uint(mask)))

    def kernel_stats(self, reset: bool = True):
        L = load_library()
        out = {}
        for kid, name in enumerate(KERNELS):
            n = ct.c_int64()
            ms = ct.c_double()
            by = ct.c_double()
            _check(L.cgs_kernel_stats(self._h, kid, ct.byref(n), ct.byref(ms), ct.byref(by), int(reset)))
            out[name] = {"launches": n.value, "ms": ms.value, "bytes": by.value}
        return out


# ---------------------------------------------------------------------------
# codeare-shaped operator classes (Params dict in, same keys and defaults)

def _imsz(p):
    return tuple(int(v) for v in p["imsz"])


class NFFT:
    """NFFT<cxfl> (src/matrix/ft/NFFT.hpp): keys nk, imsz, m (1), alpha (1.0),
    epsilon (7e-4), maxit (3), 3rd_dim_cart (False), b0 and timing (the
    off-resonance model, NFFT.hpp:117-173: b0 (N0, N1) in cycles per unit of
    timing, timing (nk,); applied once KSpace is set).  ``*`` is Trafo,
    ``adjoint`` (->*) the iterative NFFT3 inverse; ``adjoint_true`` is
    nfft_adjoint.  With 3rd_dim_cart an imsz (N0, N1, Nz) is Nz 2D slices on
    the one trajectory (NFFT.hpp:84-103): Trafo maps (N0, N1, Nz) to (nk, Nz)
    slice by slice and Adjoint back (NFFT.hpp:331-360, :380-412; codeare leaves
    the Cartesian FFT along the slices commented out, and so does this class)."""

    def __init__(self, params: dict):
        if "nk" not in params or "imsz" not in params:
            raise ValueError("NFFT: nk and imsz must be given (NFFT.hpp:72-98)")
        self.M = int(params["nk"])
        self.N = _imsz(params)
        self.ncart = 1
        if bool(params.get("3rd_dim_cart", False)) and len(self.N) == 3:
            self.ncart = self.N[-1]
            self.N = self.N[:-1]
        self.m = int(params.get("m", 1))
        self.alpha = float(params.get("alpha", 1.0))
        self.epsilon = float(params.get("epsilon", 7.0e-4))
        self.maxit = int(params.get("maxit", 3))
        # the slices are the plan's image batch
        self.plan = Plan(self.N, self.M, self.alpha, self.m, ncoils=self.ncart)
        self._b0 = None
        if params.get("b0") is not None:
            b0 = np.asarray(params["b0"], dtype=np.float32)
            t = params.get("timing")
            # NFFT.hpp:117-141: a b0 that does not fit the image, or timing that
            # does not fit the trajectory, falls back to b0 = 0 with a warning
            if b0.size != int(np.prod(self.N)) or t is None or np.asarray(t).size != self.M:
                import warnings
                warnings.warn("NFFT: b0 / timing do not fit image / trajectory; defaulting to b0 = 0")
            else:
                self._b0 = (b0, np.asarray(t, dtype=np.float32))

    def KSpace(self, k):
        self.plan.kspace(k)
        if self._b0 is not None:
            self.plan.b0(*self._b0)

    def Weights(self, w):
        self.plan.weights(w)

    def Trafo(self, img):
        return self.plan.nfft_forward(img)

    __mul__ = Trafo

    def Adjoint(self, f):
        return self.plan.nfft_inverse_ref(f, maxit=self.maxit, eps=self.epsilon)

    def adjoint_true(self, f):
        return self.plan.nfft_adjoint(f)


class NCSENSE:
    """NCSENSE<cxfl> (src/matrix/ft/NCSENSE.hpp:69-184): keys sensitivities, nk,
    ftiter (3), cgiter (0), cgeps (1e-6), lambda (1e-6), m (1), alpha (1.0),
    fteps (1e-3), dim4/dim5 (frames, with frame_offsets for uneven counts),
    3rd_dim_cart (False: with True the maps are (N0, N1, Nz, C), the data
    (nk, Nz, C), and the Nz slices share the trajectory and the weights,
    NCSENSE.hpp:81-99, :338-339).

    ``*`` = E, ``/`` = reference E^H (ITrafo per coil, intensity-corrected),
    ``Adjoint`` = CGLS solve (codeare semantics: squared-ratio cgeps stop, the
    reference E^H, one CG over all frames or slices as CGLS runs on the whole
    image matrix).  ``mode="true"`` switches E^H to the density-weighted true
    adjoint for both ``/`` and the solve; ``joint=False`` gives every frame its
    own CG."""

    def __init__(self, params: dict, mode: str = "ref", joint: bool = True):
        S = _c64(params["sensitivities"])
        self.cart = bool(params.get("3rd_dim_cart", False))
        if self.cart and S.ndim != 4:
            raise ValueError("NCSENSE: 3rd_dim_cart needs (N0, N1, Nz, C) maps")
        self.d = S.ndim - (2 if self.cart else 1)
        self.C = S.shape[-1]
        self.N = tuple(S.shape[: self.d])
        self.M = int(params.get("nk", 1))
        self.ftiter = int(params.get("ftiter", 3))
        self.fteps = float(params.get("fteps", 1.0e-3))
        self.cgiter = int(params.get("cgiter", 0))
        self.cgeps = float(params.get("cgeps", 1.0e-6))
        self.lam = float(params.get("lambda", 1.0e-6))
        self.m = int(params.get("m", 1))
        self.alpha = float(params.get("alpha", 1.0))
        self.mode = CGS_EH_REF if mode == "ref" else CGS_EH_TRUE
        self.joint = bool(joint)
        frames = S.shape[2] if self.cart else int(params.get("dim4", 1)) * int(params.get("dim5", 1))
        foff = params.get("frame_offsets")
        if foff is None or self.cart:
            foff = np.arange(frames + 1, dtype=np.int64) * self.M
        foff = np.asarray(foff, dtype=np.int64)
        self.F = len(foff) - 1
        self.plan = Plan(self.N, int(foff[-1]), self.alpha, self.m, ncoils=self.C,
                         frame_offsets=foff, maps_per_frame=self.cart)
        self.plan.sensitivities(S)

    def Sensitivities(self, S):
        self.plan.sensitivities(S)

    def _per_frame(self, v, block, per_sample):
        # one block shared by every frame/slice (NCSENSE.hpp:204-248) -> one per
        # frame; per-frame blocks (uneven ones with frame_offsets) pass through
        v = np.asarray(v, dtype=np.float32)
        if v.size == block * self.F or v.size == per_sample * self.plan.M or self.F == 1:
            return v
        if v.size != block:
            raise ValueError("k-space / weights size does not fit the frames")
        return np.concatenate([v.reshape(-1, order="F")] * self.F)

    def KSpace(self, k):
        k = self._per_frame(k, self.d * self.M, self.d)
        self.plan.kspace(np.asarray(k, dtype=np.float32).reshape((self.d, -1), order="F"))

    def Weights(self, w):
        self.plan.weights(self._per_frame(w, self.M, 1))

    # codeare's slice data (nk, Nz, C) <-> the plan's frame blocks (nk, C) per slice
    def _to_plan(self, y):
        y = _c64(y)
        if not self.cart:
            return y
        return _c64(np.transpose(y.reshape((self.M, self.F, self.C), order="F"), (0, 2, 1)))

    def Trafo(self, x):
        y = self.plan.E(x)
        if not self.cart:
            return y
        return np.asfortranarray(np.transpose(y.reshape((self.M, self.C, self.F), order="F"), (0, 2, 1)))

    __mul__ = Trafo

    def __truediv__(self, y):
        return self.plan.EH(self._to_plan(y), self.mode, self.ftiter, self.fteps)

    def Adjoint(self, y, x0=None):
        x, self.last_report = self.plan.solve(
            self._to_plan(y), x0=x0, maxit=self.cgiter, tol=self.cgeps, tol_kind=CGS_TOL_SQUARED_RATIO,
            lam=self.lam, eh_mode=self.mode, ftiter=self.ftiter, fteps=self.fteps, joint=self.joint)
        return x
