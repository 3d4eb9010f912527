"""codeare's reconstruction modules for this path, as host-side strategies over
the GPU operators: `CGSENSE` (src/modules/CGSENSE.cpp:52-178), `NuFFT`
(src/modules/NuFFT.cpp:49-142) and `EstimateSensitivities`
(src/modules/EstimateSensitivities.cpp:49-172).

A strategy follows codeare's ReconStrategy life cycle (Init / Prepare /
Process, src/core/ReconStrategy.hpp) over a workspace of named matrices (a
dict standing in for the Workspace singleton, Workspace.hpp:59-282).  The XML
attributes keep codeare's names, and the quirks of the modules are kept:
CGSENSE passes neither `alpha` nor `fteps` to NCSENSE (so sigma = 1 and fteps =
1e-3), and NuFFT reads `maxit`/`epsilon` but does not pass them to NFFT (so
the iterative adjoint runs NFFT's defaults, 3 steps and 7e-4).

Command line (files in codeare's .cod format):

    python -m codeare_amd.modules CGSENSE nk=16384 m=3 lambda=5e-2 cgeps=1e-7 \
        cgmaxit=10 ftmaxit=3 sensitivities=s.cod kspace=k.cod weights=w.cod \
        signals=y.cod image=out.cod
"""
from __future__ import annotations

import sys

import numpy as np

from .cgsense import NCSENSE, NFFT
from .codfile import read_cod, write_cod


# Shepp-Logan ellipses as codeare's phantom() tabulates them (Creators.hpp:454-489):
# semi-axes (a0, a1) and centre offsets (c0, c1) in half-widths, angle, value
_SL2 = [(.69, .92, .0, .0, .0, 1.), (.6624, .874, .0, -.0184, .0, -.8), (.11, .31, -.22, .0, -.3, -.2),
        (.16, .41, .22, .0, .3, -.2), (.21, .25, .0, .35, .0, .1), (.046, .046, .00, .1, .0, .1),
        (.046, .046, .0, -.1, .0, .1), (.046, .023, .08, -.605, .0, .1), (.023, .023, .0, -.606, .0, .1),
        (.023, .046, -.06, -.605, .0, .1)]
# Koay's 3D ellipsoids (Creators.hpp:510-533): semi-axes, centre offsets, angle
_SL3 = [(.69, .92, .9, .0, .0, .0, .0), (.662, .874, .88, .0, .0, .0, .0), (.11, .31, .22, -.22, .0, -.25, -.3),
        (.16, .41, .21, .22, .0, -.25, .3), (.21, .25, .5, .0, .35, -.25, .0), (.046, .046, .046, .0, .1, -.25, .0),
        (.046, .023, .02, .08, -.65, -.25, .0), (.046, .023, .02, .06, -.65, -.25, .0),
        (.056, .04, .1, -.06, -.105, .625, .0), (.056, .056, .1, .0, .1, .625, .0)]


def phantom(n: int, d: int = 2) -> np.ndarray:
    """codeare's phantom<cxfl>(n) (2D Shepp-Logan, Creators.hpp:369-489) or
    phantom3D (Creators.hpp:408-536) as a complex64 column-major array.  The
    ellipse test runs in FP32 as codeare's does: element (c, r) lies inside when
    x^2 + y^2 <= 1 with x = ((c - m1) cos + (r - m0) sin) / a1 and y = ((r - m0)
    cos - (c - m1) sin) / a0, m = (1 - centre) n / 2, a = semi-axis n / 2.  In 3D
    codeare's ellipsoid writes the slice index s where its value belongs (its
    loop variable shadows the value parameter, Creators.hpp:427-433); that is
    kept, so the 3D phantom holds slice indices inside its ellipsoids."""
    f = np.float32
    n_ = f(n)
    c = np.arange(n, dtype=f)
    if d == 2:
        res = np.zeros((n, n), dtype=np.complex64)
        C, R = np.meshgrid(c, c, indexing="ij")           # res(c, r): first index c
        for a0, a1, p2, p3, ang, val in _SL2:
            A0, A1 = f(a0) * n_ / f(2), f(a1) * n_ / f(2)
            M0, M1 = (f(1) - f(p2)) * n_ / f(2), (f(1) - f(p3)) * n_ / f(2)
            cs, sn = f(np.cos(f(ang))), f(np.sin(f(ang)))
            x = ((C - M1) * cs + (R - M0) * sn) / A1
            y = ((R - M0) * cs - (C - M1) * sn) / A0
            res += np.where(x * x + y * y <= f(1), np.complex64(val), np.complex64(0))
        return np.asfortranarray(res)
    res = np.zeros((n, n, n), dtype=np.complex64)
    C, R, S = np.meshgrid(c, c, c, indexing="ij")          # res(c, r, s)
    for a0, a1, a2, p3, p4, p5, ang in _SL3:
        A0, A1, A2 = f(a0) * n_ / f(2), f(a1) * n_ / f(2), f(a2) * n_ / f(2)
        M0, M1, M2 = (f(1) - f(p3)) * n_ / f(2), (f(1) - f(p4)) * n_ / f(2), (f(1) - f(p5)) * n_ / f(2)
        cs, sn = f(np.cos(f(ang))), f(np.sin(f(ang)))
        x = ((C - M1) * cs + (R - M0) * sn) / A1
        y = ((R - M0) * cs - (C - M1) * sn) / A0
        z = (S - M2) / A2
        res += np.where(x * x + y * y + z * z <= f(1), S.astype(np.complex64), np.complex64(0))
    return np.asfortranarray(res)


class ReconStrategy:
    """Init(attributes) / Prepare(workspace) / Process(workspace)."""

    def __init__(self):
        self.attrs = {}
        self.initialised = False

    def attribute(self, name, default=None, cast=float):
        v = self.attrs.get(name, default)
        return None if v is None else cast(v)


class CGSENSE(ReconStrategy):
    """Non-Cartesian SENSE: workspace "sensitivities" (N..., C) complex,
    "kspace" (d, nk) float, "weights" (nk) float, "signals" (nk, C) complex
    -> "image" (N...) complex (CGSENSE.cpp:132-178).  With dft_3rd_dim the
    maps are (N0, N1, Nz, C), the signals (nk, Nz, C) and the image
    (N0, N1, Nz): stack-of-stars slices on one 2D trajectory, one CG."""

    def init(self, **attrs):
        # defaults of CGSENSE::Init (CGSENSE.cpp:61-64) and of the XML examples
        self.attrs = {"testcase": 0, "verbose": 0, "noise": 0.0, "lambda": 5.0e-2,
                      "cgeps": 1.0e-7, "cgmaxit": 10, "ftmaxit": 3, "m": 1, "dft_3rd_dim": 0}
        self.attrs.update(attrs)
        if "nk" not in self.attrs:
            raise KeyError("CGSENSE needs the nk attribute")
        self.initialised = True

    def prepare(self, ws: dict):
        a = self.attribute
        params = {"sensitivities": ws["sensitivities"], "nk": a("nk", cast=int),
                  "ftiter": a("ftmaxit", cast=int), "cgiter": a("cgmaxit", cast=int),
                  "cgeps": a("cgeps"), "lambda": a("lambda"), "m": a("m", cast=int),
                  "3rd_dim_cart": bool(a("dft_3rd_dim", cast=int))}   # CGSENSE.cpp:147
        # no "alpha" and no "fteps": NCSENSE's defaults (sigma 1, 1e-3) apply
        self.ncs = NCSENSE(params)
        self.ncs.KSpace(ws.pop("kspace"))
        self.ncs.Weights(ws.pop("weights"))

    def process(self, ws: dict):
        if int(self.attrs["testcase"]):
            # E applied to codeare's Shepp-Logan phantom of the maps' side (CGSENSE.cpp:171-172)
            S = np.asarray(ws["sensitivities"])
            data = self.ncs * phantom(S.shape[0], 2)
        else:
            data = ws["signals"]
        noise = float(self.attrs["noise"])
        if noise:
            rng = np.random.default_rng()
            data = data + noise * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))
        ws["image"] = self.ncs.Adjoint(np.asarray(data, dtype=np.complex64))
        return ws["image"]


class NuFFT(ReconStrategy):
    """Single-coil NFFT: workspace "kspace", "weights", "data" (M*shots)
    -> "img" = NFFT ->* data (NuFFT.cpp:49-142).  With dft_3rd_dim (dim 3) the
    data is (M*shots, Nz) and the image (Nx, Ny, Nz), one 2D slice per column."""

    def init(self, **attrs):
        self.attrs = {"m": 1, "alpha": 1.0, "shots": 1, "test_case": 0, "dft_3rd_dim": 0}
        self.attrs.update(attrs)
        d = int(self.attrs["dim"])
        sides = ["Nx", "Ny", "Nz"][:d]
        imsz = [int(self.attrs.get(s, 1)) for s in sides]
        nk = int(self.attrs["M"]) * int(self.attrs["shots"])
        # maxit and epsilon are read but not passed on (NuFFT.cpp:80-107)
        self.ft = NFFT({"alpha": float(self.attrs["alpha"]), "m": int(self.attrs["m"]),
                        "nk": nk, "imsz": imsz, "3rd_dim_cart": bool(int(self.attrs["dft_3rd_dim"]))})
        self.initialised = True

    def prepare(self, ws: dict):
        self.ft.KSpace(ws.pop("kspace"))
        self.ft.Weights(ws.pop("weights"))

    def process(self, ws: dict):
        if int(self.attrs["test_case"]):
            # NFFT applied to codeare's phantom of the NFFT's image side, 3D for dim 3
            # (NuFFT.cpp:138-140; ImageSize() is N[0], NFFT.hpp:445)
            d = int(self.attrs["dim"])
            data = self.ft * phantom(self.ft.N[0], 3 if d == 3 else 2)
        else:
            data = ws["data"]
        ws["img"] = self.ft.Adjoint(np.asarray(data, dtype=np.complex64).reshape(-1, order="F"))
        return ws["img"]


NYUGA = 111.246117975   # the golden angle in degrees (EstimateSensitivities.cpp:29-31)


def natural_cspline(x, y, xi):
    """Natural cubic spline of the columns of y (len(x) rows) at xi -- the
    interpolant of GSL's cspline type, codeare's interp1 default
    (Interpolate.hpp:18-38, PolyVal.hpp:261-274): second derivatives zero at
    both ends, continuous first and second derivatives at the knots; xi within
    [x[0], x[-1]].  Solved in FP64 with the tridiagonal (Thomas) recurrence."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    y2 = y.reshape(len(x), -1)
    n = len(x)
    xi = np.asarray(xi, np.float64)
    if n < 3:   # GSL needs 3 points for a cspline; fall back to the line
        out = np.stack([np.interp(xi, x, y2[:, j]) for j in range(y2.shape[1])], axis=1)
        return out.reshape((len(xi),) + y.shape[1:])
    h = np.diff(x)
    # interior second derivatives M_1..M_{n-2}:
    # h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (d_i - d_{i-1})
    d = np.diff(y2, axis=0) / h[:, None]
    rhs = 6.0 * (d[1:] - d[:-1])
    a = h[:-1].copy()              # sub-diagonal
    b = 2.0 * (h[:-1] + h[1:])     # diagonal
    c = h[1:].copy()               # super-diagonal
    k = len(b)
    cp = np.zeros(k)
    dp = np.zeros_like(rhs)
    cp[0] = c[0] / b[0]
    dp[0] = rhs[0] / b[0]
    for i in range(1, k):
        den = b[i] - a[i] * cp[i - 1]
        cp[i] = c[i] / den if i < k - 1 else 0.0
        dp[i] = (rhs[i] - a[i] * dp[i - 1]) / den
    Mi = np.zeros_like(rhs)
    Mi[-1] = dp[-1]
    for i in range(k - 2, -1, -1):
        Mi[i] = dp[i] - cp[i] * Mi[i + 1]
    M = np.zeros_like(y2)
    M[1:-1] = Mi
    j = np.clip(np.searchsorted(x, xi, side="right") - 1, 0, n - 2)
    t = xi - x[j]
    hj = h[j][:, None]
    yl, yr, Ml, Mr = y2[j], y2[j + 1], M[j], M[j + 1]
    out = (yl + t[:, None] * ((yr - yl) / hj - hj * (2.0 * Ml + Mr) / 6.0)
           + t[:, None] ** 2 * Ml / 2.0 + t[:, None] ** 3 * (Mr - Ml) / (6.0 * hj))
    return out.reshape((len(xi),) + y.shape[1:])


class EstimateSensitivities(ReconStrategy):
    """Coil images from multi-channel golden-angle radial data
    (src/modules/EstimateSensitivities.cpp:49-172): workspace "meas" (nk, nc,
    nl, nv) complex and "sync" (the gradient-raster trajectory, (kx, ky)
    interleaved in its first column) -> "sensitivities" (image_size..., nc)
    complex and "kspace" (2, nk * nv).

    Steps as the module takes them: meas *= 1e6; the raster trajectory is
    interpolated 10x finer with interp1's natural cubic spline and cut to nk
    points; volume v is that spoke rotated by v golden angles; the spokes are
    scaled to max |k| = 1/2; the NFFT (ftalpha, ftm, ftiter; unit weights)
    inverts the coil's data with its iterative adjoint (ITrafo).  Kept quirks:
    only coil 0 is reconstructed (the coil loop runs `i < 1`, :147-150; the
    other coils stay zero unless all_coils=1), and the sum of squares is formed
    but neither divides nor normalises the maps (:153-164), so neither is done.
    One departure: the module writes the rotation as cpolar(1, exp(v * NYUGA))
    (:123), whose angle overflows FP32 for every v >= 1 and makes those spokes
    NaN; the rotation here is the golden angle the constant names, v * NYUGA
    degrees."""

    def init(self, **attrs):
        self.attrs = {"ftalpha": 1.0, "ftm": 1, "ftiter": 3, "cart_3rd_dim": 0, "all_coils": 0}
        self.attrs.update(attrs)
        for key in ("nk", "nl", "image_size"):
            if key not in self.attrs:
                raise KeyError(f"EstimateSensitivities needs the {key} attribute")
        imsz = self.attrs["image_size"]
        if isinstance(imsz, str):
            imsz = [int(v) for v in imsz.replace(",", " ").split()]
        self.imsz = [int(v) for v in imsz]
        self.nk_attr = int(self.attrs["nk"]) * int(self.attrs["nl"])
        self.ft = NFFT({"alpha": float(self.attrs["ftalpha"]), "m": int(self.attrs["ftm"]),
                        "nk": self.nk_attr, "imsz": self.imsz, "maxit": int(self.attrs["ftiter"]),
                        "3rd_dim_cart": bool(int(self.attrs["cart_3rd_dim"]))})
        self.initialised = True

    def prepare(self, ws: dict):
        sync = np.asarray(ws["sync"], dtype=np.float32)
        col = sync.reshape(sync.shape[0], -1, order="F")[:, 0]     # kspace(CR(), CR(0)), squeezed
        self.raster = col.reshape(2, -1, order="F").T.copy()        # resize to (2, n/2), permute: (n/2, 2)

    def trajectory(self, nk: int, nv: int) -> np.ndarray:
        """(2, nk * nv) float32: the interpolated spoke, rotated by v golden
        angles for volume v, scaled to max |k| = 1/2."""
        r = self.raster.astype(np.float32)
        nr = r.shape[0]
        xi = np.linspace(0.0, 1.0, nr, dtype=np.float32)
        XI = np.linspace(0.0, 1.0, nr * 10, dtype=np.float32)
        ki = natural_cspline(xi, r, XI).astype(np.float32)[:nk]
        spoke = (ki[:, 0] + 1j * ki[:, 1]).astype(np.complex64)
        ang = np.deg2rad(np.arange(nv, dtype=np.float64) * NYUGA)
        kt = (spoke[:, None] * np.exp(1j * ang)[None, :].astype(np.complex64)).astype(np.complex64)
        kt /= np.float32(2.0) * np.max(np.abs(kt))
        kt = kt.reshape(-1, order="F")
        return np.asfortranarray(np.stack([kt.real, kt.imag]).astype(np.float32))

    def process(self, ws: dict):
        data = np.asarray(ws["meas"], dtype=np.complex64) * np.float32(1e6)
        data = np.squeeze(data)
        while data.ndim < 4:
            data = data[..., None]
        nk, nc, nl, nv = data.shape[:4]
        cart = bool(int(self.attrs["cart_3rd_dim"]))
        if nk * nv != self.nk_attr or nl != (self.imsz[-1] if cart else 1):
            raise ValueError(f"meas holds {nk} samples x {nv} volumes x {nl} lines per coil; the NFFT was made "
                             f"for nk * nl = {self.nk_attr} samples" + (" per slice" if cart else ""))
        k = self.trajectory(nk, nv)
        self.ft.KSpace(k)
        self.ft.Weights(np.ones(nk * nv, np.float32))
        tmp = np.asfortranarray(np.transpose(data, (0, 3, 2, 1)))   # (nk, nv, nl, nc)
        sens = np.zeros(tuple(self.imsz) + (nc,), dtype=np.complex64, order="F")
        for i in range(nc if int(self.attrs["all_coils"]) else 1):
            f = tmp[..., i].reshape(-1, order="F")
            if cart:
                f = f.reshape((nk * nv, nl), order="F")
            sens[..., i] = np.asarray(self.ft.Adjoint(f)).reshape(self.imsz, order="F")
        ws["sensitivities"] = sens
        ws["kspace"] = k
        return sens


STRATEGIES = {"CGSENSE": (CGSENSE, ("sensitivities", "kspace", "weights", "signals"), "image"),
              "NuFFT": (NuFFT, ("kspace", "weights", "data"), "img"),
              "EstimateSensitivities": (EstimateSensitivities, ("meas", "sync"), "sensitivities")}


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in STRATEGIES:
        print(__doc__)
        return 2
    cls, inputs, output = STRATEGIES[argv[0]]
    kv = dict(a.split("=", 1) for a in argv[1:])
    ws = {name: read_cod(kv.pop(name))[0] for name in inputs if name in kv}
    out_path = kv.pop(output, None)
    s = cls()
    s.init(**kv)
    s.prepare(ws)
    img = s.process(ws)
    if out_path:
        write_cod(out_path, np.asarray(img, dtype=np.complex64), output)
    return 0


if __name__ == "__main__":
    sys.exit(main())
