"""Seeded synthetic inputs for the non-Cartesian CG-SENSE workloads.

Everything is generated on the host with numpy and is identical for the GPU
path, the CPU oracle and the benchmark.  Arrays use codeare's column-major
conventions (Matrix.hpp:579): images (N0, N1[, N2]) in Fortran order, coil maps
(N0, N1[, N2], C), trajectories (d, M) float32 with k in [-0.5, 0.5), data (M, C).

Coordinates: the field of view is [-0.5, 0.5)^d and pixel i sits at (i - N/2)/N.

Workloads (sizes fixed by the CG-SENSE benchmark definitions):
  1  128^2,  8 coils,  64 uniform radial spokes x 256, sigma 2, m 3, 10 CG its
  2  256^2, 32 coils, 402 golden-angle spokes x 512, sigma 2, m 3, to 1e-5, lambda 1e-3
  3  128^3, 16 coils, 20000 Fibonacci-sphere spokes x 256, sigma 1.5, m 3, 15 CG its
  4  64 frames of 192^2, 16 coils, 13..89 golden-angle spokes x 384 per frame
  5  workload 2 shapes, many seeded phantoms (learned warm start)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GOLDEN_ANGLE_DEG = 111.2461
SEED_BASE = 20261202


def seed_for(cfg: int, item: int = 0) -> int:
    return SEED_BASE + 1000 * int(cfg) + int(item)


@dataclass
class Workload:
    name: str
    d: int
    N: int
    coils: int
    samples: int              # readout samples per spoke
    sigma: float = 2.0
    m: int = 3
    ellipses: int = 10
    density: str = "k"        # "k": max(|k|, 1/R); "k2": max(|k|^2, 1/R^2)
    lam: float = 0.0
    maxit: int = 10
    tol: float = 0.0          # 0 = fixed iteration count
    snr_db: float = 40.0
    spokes: int = 0
    golden: bool = True
    frames: int = 1
    spoke_range: tuple = field(default_factory=tuple)


WORKLOADS = {
    1: Workload("radial2d_128_c8", 2, 128, 8, 256, 2.0, 3, 10, "k", 0.0, 10, 0.0,
                spokes=64, golden=False),
    2: Workload("golden2d_256_c32", 2, 256, 32, 512, 2.0, 3, 20, "k", 1e-3, 100, 1e-5,
                spokes=402, golden=True),
    3: Workload("koosh3d_128_c16", 3, 128, 16, 256, 1.5, 3, 15, "k2", 0.0, 15, 0.0,
                spokes=20000),
    4: Workload("dynamic2d_192_c16_f64", 2, 192, 16, 384, 2.0, 3, 20, "k", 0.0, 100, 1e-5,
                golden=True, frames=64, spoke_range=(13, 89)),
    5: Workload("golden2d_256_c32_warm", 2, 256, 32, 512, 2.0, 3, 20, "k", 1e-3, 100, 1e-5,
                spokes=402, golden=True),
}


# ---------------------------------------------------------------------------
def pixel_coords(N: int, d: int):
    """Per-axis pixel positions (i - N/2)/N, as an 'ij' meshgrid."""
    r = (np.arange(N) - N // 2) / N
    return np.meshgrid(*([r] * d), indexing="ij")


def _rotation3(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d],
    ])


def ellipse_params(rng, n_ell: int, d: int):
    out = []
    for _ in range(n_ell):
        c = rng.uniform(-0.6, 0.6, size=d)
        ax = rng.uniform(0.05, 0.4, size=d)
        if d == 2:
            th = rng.uniform(0, math.pi)
            R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        else:
            R = _rotation3(rng)
        v = rng.uniform(0.1, 1.0)
        ph = rng.uniform(0, 2 * math.pi)    # used only by the dynamic series
        out.append((c, ax, R, v, ph))
    return out


def phantom(N: int, d: int, params, frame: int | None = None, nframes: int = 64) -> np.ndarray:
    """Sum of ellipses/ellipsoids; centres and semi-axes in FOV-half units."""
    pos = [2.0 * g for g in pixel_coords(N, d)]          # FOV-half units
    img = np.zeros((N,) * d, dtype=np.float64)
    for c, ax, R, v, ph in params:
        rel = [pos[a] - c[a] for a in range(d)]
        # coordinates in the ellipse frame: R^T (p - c)
        acc = np.zeros_like(img)
        for i in range(d):
            u = sum(R[a, i] * rel[a] for a in range(d))
            acc += (u / ax[i]) ** 2
        val = v
        if frame is not None:
            val = v * (1.0 + 0.5 * math.sin(2 * math.pi * frame / nframes + ph))
        img += np.where(acc <= 1.0, val, 0.0)
    return np.asfortranarray(img.astype(np.complex64))


def fibonacci_sphere(S: int) -> np.ndarray:
    s = np.arange(S)
    z = 1.0 - (2 * s + 1) / S
    phi = s * math.pi * (3.0 - math.sqrt(5.0))
    rr = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([rr * np.cos(phi), rr * np.sin(phi), z])


def coil_maps(N: int, d: int, C: int, rng) -> np.ndarray:
    """Gaussian coils (sigma 0.5 FOV) on a ring/sphere of radius 0.7 FOV with a
    random linear phase exp(i g.r), g ~ U(-pi, pi)^d per FOV; sum |S|^2 = 1."""
    pos = pixel_coords(N, d)
    if d == 2:
        ang = 2 * math.pi * np.arange(C) / C
        centres = 0.7 * np.stack([np.cos(ang), np.sin(ang)])
    else:
        centres = 0.7 * fibonacci_sphere(C)
    S = np.zeros((N,) * d + (C,), dtype=np.complex128)
    for c in range(C):
        r2 = sum((pos[a] - centres[a, c]) ** 2 for a in range(d))
        g = rng.uniform(-math.pi, math.pi, size=d)
        ph = sum(g[a] * pos[a] for a in range(d))
        S[..., c] = np.exp(-r2 / (2 * 0.5 ** 2)) * np.exp(1j * ph)
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=-1, keepdims=True))
    return np.asfortranarray(S.astype(np.complex64))


def radial_2d(spokes: int, R: int, golden: bool, first_spoke: int = 0) -> np.ndarray:
    s = np.arange(first_spoke, first_spoke + spokes)
    if golden:
        th = np.deg2rad(np.mod(s * GOLDEN_ANGLE_DEG, 360.0))
    else:
        th = math.pi * s / spokes
    kr = (np.arange(R) - R // 2) / R
    k0 = (np.cos(th)[:, None] * kr[None, :]).reshape(-1)
    k1 = (np.sin(th)[:, None] * kr[None, :]).reshape(-1)
    return np.asfortranarray(np.stack([k0, k1]).astype(np.float32))


def koosh_3d(spokes: int, R: int) -> np.ndarray:
    dirs = fibonacci_sphere(spokes)
    kr = (np.arange(R) - R // 2) / R
    k = dirs[:, :, None] * kr[None, None, :]
    return np.asfortranarray(k.reshape(3, -1).astype(np.float32))


def density(k: np.ndarray, kind: str, R: int) -> np.ndarray:
    kk = np.sqrt(np.sum(k.astype(np.float64) ** 2, axis=0))
    if kind == "k2":
        w = np.maximum(kk ** 2, 1.0 / R ** 2)
    else:
        w = np.maximum(kk, 1.0 / R)
    return w.astype(np.float32)


def add_noise(y: np.ndarray, snr_db: float, rng) -> np.ndarray:
    rms = math.sqrt(float(np.mean(np.abs(y.astype(np.complex128)) ** 2)))
    sd = rms * 10 ** (-snr_db / 20.0) / math.sqrt(2.0)
    n = rng.normal(0, sd, size=y.shape) + 1j * rng.normal(0, sd, size=y.shape)
    return np.asfortranarray((y + n).astype(np.complex64))


# ---------------------------------------------------------------------------
@dataclass
class Problem:
    wl: Workload
    image: np.ndarray        # (N,)*d or (N,)*d + (F,)
    maps: np.ndarray         # (N,)*d + (C,)
    k: np.ndarray            # (d, M_total)
    w: np.ndarray            # (M_total,)
    frame_offsets: np.ndarray  # (F+1,) int64
    rng_state: object = None


def make_problem(cfg: int, item: int = 0, *, N: int | None = None, coils: int | None = None,
                 spokes: int | None = None, samples: int | None = None,
                 frames: int | None = None) -> Problem:
    """Phantom, maps, trajectory and density weights of a workload (no data yet:
    data needs the forward operator, see ``make_data``).  Size overrides give
    reduced problems of the same construction for tests."""
    base = WORKLOADS[cfg]
    wl = Workload(**{**base.__dict__})
    if N is not None:
        wl.N = N
    if coils is not None:
        wl.coils = coils
    if spokes is not None:
        wl.spokes = spokes
    if samples is not None:
        wl.samples = samples
    if frames is not None:
        wl.frames = frames
    rng = np.random.default_rng(seed_for(cfg, item))
    d, Nn, C, R = wl.d, wl.N, wl.coils, wl.samples
    params = ellipse_params(rng, wl.ellipses, d)
    maps = coil_maps(Nn, d, C, rng)
    if wl.frames > 1:
        lo, hi = wl.spoke_range if wl.spoke_range else (wl.spokes, wl.spokes)
        counts = rng.integers(lo, hi + 1, size=wl.frames)
        ks, first = [], 0
        for f in range(wl.frames):
            ks.append(radial_2d(int(counts[f]), R, True, first))
            first += int(counts[f])
        k = np.asfortranarray(np.concatenate(ks, axis=1))
        off = np.concatenate([[0], np.cumsum(counts * R)]).astype(np.int64)
        img = np.stack([phantom(Nn, d, params, f, wl.frames) for f in range(wl.frames)], axis=-1)
        img = np.asfortranarray(img)
    else:
        if d == 2:
            k = radial_2d(wl.spokes, R, wl.golden)
        else:
            k = koosh_3d(wl.spokes, R)
        off = np.array([0, k.shape[1]], dtype=np.int64)
        img = phantom(Nn, d, params)
    w = density(k, wl.density, R)
    return Problem(wl, img, maps, k, w, off, rng)


def make_data(prob: Problem, E_fn) -> np.ndarray:
    """y = E x + complex Gaussian noise at the workload's SNR, with E_fn(x)
    the forward encoding operator (GPU or oracle).  Noise uses the problem's
    own seeded stream, so the same problem always gets the same noise."""
    y = E_fn(prob.image)
    return add_noise(y, prob.wl.snr_db, prob.rng_state)
