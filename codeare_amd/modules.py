"""codeare's reconstruction modules for this path, as host-side strategies over
the GPU operators: `CGSENSE` (src/modules/CGSENSE.cpp:52-178) and `NuFFT`
(src/modules/NuFFT.cpp:49-142).

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
            raise NotImplementedError("the built-in phantom test case is not built; pass signals")
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
            raise NotImplementedError("the built-in phantom test case is not built; pass data")
        ws["img"] = self.ft.Adjoint(np.asarray(ws["data"], dtype=np.complex64).reshape(-1, order="F"))
        return ws["img"]


STRATEGIES = {"CGSENSE": (CGSENSE, ("sensitivities", "kspace", "weights", "signals"), "image"),
              "NuFFT": (NuFFT, ("kspace", "weights", "data"), "img")}


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
