"""Regression tests of fixed defects in the solver's host logic (GPU)."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _plan(cgs, N, coils, spokes, samples):
    from codeare_amd import data
    prob = data.make_problem(2, 0, N=N, coils=coils, spokes=spokes, samples=samples)
    wl = prob.wl
    M = prob.k.shape[1]
    P = cgs.Plan((N, N), M, wl.sigma, wl.m, ncoils=coils)
    P.kspace(prob.k)
    P.weights(prob.w)
    P.sensitivities(prob.maps)
    return prob, P


@pytest.mark.parametrize("N,coils", [(128, 16), (64, 8)])   # lean Toeplitz passes / generic passes
def test_graph_solve_after_history_regrowth(cgs, N, coils):
    """A captured CG graph writes the residual history through the buffer it
    was captured with.  A later solve with a larger maxit regrows that buffer;
    the next graph solve with the first options must not replay the old graph
    into the freed buffer: its iterations, history and image must equal a
    fresh plan's."""
    from codeare_amd import data
    prob, P = _plan(cgs, N, coils, 96, 2 * N)
    y = data.make_data(prob, P.E)
    opts = dict(maxit=100, tol=1e-5, lam=1e-3, toeplitz=True, graph_iters=8)
    x1, r1 = P.solve(y, **opts)
    h1 = P.residual_history()
    P.solve(y, maxit=3000, tol=1e-5, lam=1e-3, toeplitz=True, graph_iters=0)   # regrows the history
    x2, r2 = P.solve(y, **opts)
    h2 = P.residual_history()
    _, Q = _plan(cgs, N, coils, 96, 2 * N)
    x3, r3 = Q.solve(y, **opts)
    h3 = Q.residual_history()
    print(f"N={N} C={coils}: iterations {r1['iters'][0]} / {r2['iters'][0]} / fresh {r3['iters'][0]}, "
          f"history entries {len(h1)} / {len(h2)} / {len(h3)}")
    assert r1["iters"][0] == r2["iters"][0] == r3["iters"][0] == 100
    assert len(h2) == len(h3) == 101
    assert np.all(h2 > 0)
    assert np.allclose(h2, h3, rtol=1e-9, atol=0) and np.allclose(h1, h3, rtol=1e-9, atol=0)
    assert rel(x2, x3) <= 1e-6 and rel(x1, x3) <= 1e-6


def _plan_env(cgs, env, N, M, sigma, m, C, k, w, S, foff=None):
    import os
    old = {k_: os.environ.get(k_) for k_ in env}
    os.environ.update(env)
    try:
        P = cgs.Plan(N, M, sigma, m, ncoils=C, frame_offsets=foff)
    finally:
        for k_, v in old.items():
            if v is None:
                os.environ.pop(k_, None)
            else:
                os.environ[k_] = v
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    return P


@pytest.mark.parametrize("N,sigma,m,C,F,K", [
    (64, 2.0, 3, 32, 1, 256),     # 32 coils: one lane group per wave
    (64, 2.0, 3, 32, 1, 32),      # line segments split into pieces of 32 entries + reduce
    (96, 2.0, 3, 16, 2, 256),     # 16 coils: two lane groups, two frames
    (96, 2.0, 3, 16, 2, 40),      # ... with split segments
    (128, 2.0, 3, 8, 1, 256),     # 8 coils: four lane groups
    (48, 2.0, 2, 16, 1, 64),      # m = 2: 6 taps (n = 96)
    (64, 1.0, 1, 12, 1, 256),     # sigma 1, m 1 (codeare's defaults): 4 taps, 12 coils (padded lanes)
    (16, 1.5, 3, 48, 1, 24),      # n = 24 < one segment: footprints wrap inside it; 48 coils: two waves
])
def test_line_sweep_scatter_matches_tile_scatter(cgs, N, sigma, m, C, F, K):
    """The line-sweep scatter against the tile-owner scatter, on radial and
    random trajectories: E^H (weighted) and the unweighted adjoint NFFT of
    every coil, one frame or several, with line segments whole or split into
    pieces of K entries."""
    from codeare_amd import data
    from helpers import rand_img
    rng = np.random.default_rng(N + C + m)
    per = []
    for f in range(F):
        kr = data.radial_2d(24, 2 * N, True, 24 * f)
        kx = np.asfortranarray(rng.uniform(-0.5, 0.5, size=(2, 700)).astype(np.float32))
        per.append(np.concatenate([kr, kx], axis=1))
    k = np.asfortranarray(np.concatenate(per, axis=1))
    Mf = per[0].shape[1]
    M = k.shape[1]
    foff = np.arange(F + 1, dtype=np.int64) * Mf if F > 1 else None
    w = rng.uniform(0.1, 1.0, M).astype(np.float32)
    S = data.coil_maps(N, 2, C, rng)
    y = rand_img(rng, (M * C,))
    Ps = _plan_env(cgs, {"CGS_GRID_SWEEP": "1", "CGS_SWEEP_K": str(K)}, (N, N), M, sigma, m, C, k, w, S, foff)
    Pt = _plan_env(cgs, {"CGS_GRID_SWEEP": "0"}, (N, N), M, sigma, m, C, k, w, S, foff)
    e1 = rel(Ps.EH(y), Pt.EH(y))
    e2 = 0.0
    if F == 1:
        yb = y.reshape((M, C), order="F")
        e2 = rel(Ps.nfft_adjoint(yb), Pt.nfft_adjoint(yb))
    print(f"N={N} sigma={sigma} m={m} C={C} F={F} K={K}: E^H {e1:.2e}, adjoint NFFT {e2:.2e}")
    assert e1 <= 1e-6 and e2 <= 1e-6


@pytest.mark.parametrize("per_frame", [False, True])
def test_device_maps_load_matches_host_load(cgs, per_frame):
    """cgs_set_sensitivities_device (maps already on the device, loaded and the
    intensity map formed on the GPU) gives bit-identical E and E^H to the host
    loader, for shared and per-frame maps."""
    import torch
    from codeare_amd import data
    rng = np.random.default_rng(3)
    N, C, F, spokes, R = 48, 5, 3, 21, 96
    ks = [data.radial_2d(spokes, R, True, f * spokes) for f in range(F)]
    k = np.asfortranarray(np.concatenate(ks, axis=1))
    off = np.arange(F + 1, dtype=np.int64) * spokes * R
    w = np.abs(np.sqrt(k[0] ** 2 + k[1] ** 2)).astype(np.float32) + 1e-3
    nm = F if per_frame else 1
    S = np.asfortranarray(np.stack([data.coil_maps(N, 2, C, rng) for _ in range(nm)], axis=2)
                          .reshape(N, N, nm, C) if per_frame else data.coil_maps(N, 2, C, rng))
    x = torch.from_numpy((rng.normal(size=N * N * F) + 1j * rng.normal(size=N * N * F)).astype(np.complex64)).cuda()
    out = {}
    for mode in ("host", "device"):
        P = cgs.Plan((N, N), int(k.shape[1]), 2.0, 3, ncoils=C, frame_offsets=off, maps_per_frame=per_frame,
                     device_io=True)
        P.kspace(k)
        P.weights(w)
        if mode == "host":
            P.sensitivities(S)
        else:
            P.sensitivities_device(torch.from_numpy(np.asarray(S).reshape(-1, order="F")).cuda())
        y = torch.empty(int(k.shape[1]) * C, dtype=torch.complex64, device="cuda")
        xb = torch.empty_like(x)
        P.E(x, out=y)
        P.EH(y, out=xb)
        P.sync()
        out[mode] = (y.cpu().numpy(), xb.cpu().numpy())
    assert np.array_equal(out["host"][0], out["device"][0])
    assert np.array_equal(out["host"][1], out["device"][1])


@pytest.mark.parametrize("d,C,env", [
    (2, 32, {"CGS_GRID_MFMA": "0"}),                                              # register tiles
    (2, 32, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0"}),                         # batched visits
    (2, 32, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0", "CGS_GRID_ATOMIC": "1"}),  # LDS atomics
    (2, 32, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0", "CGS_GRID_BATCHED": "0"}),  # one visit at a time
    (2, 32, {"CGS_INTERP_LDS": "0"}),                                             # gather from global
    (2, 32, {"CGS_INTERP_CS": "0"}),                                              # lanes = (coil, tap row)
    (2, 12, {"CGS_INTERP_CS": "0", "CGS_INTERP_CG": "8"}),
    (2, 32, {"CGS_INTERP_CS": "1"}),                                              # lanes = (coil, sample)
    (2, 32, {"CGS_INTERP_CS": "1", "CGS_INTERP_CG": "32"}),
    (2, 12, {"CGS_INTERP_CS": "1", "CGS_INTERP_CG": "8"}),
    (2, 32, {"CGS_INTERP_CS": "2"}),                                              # ... in 512-thread blocks
    (2, 12, {"CGS_INTERP_CS": "2", "CGS_INTERP_CG": "8"}),
    (2, 32, {"CGS_INTERP_CS": "3"}),                                              # lanes = (coil pair, sample)
    (2, 12, {"CGS_INTERP_CS": "3", "CGS_INTERP_CG": "8"}),
    (2, 5, {"CGS_INTERP_CS": "3"}),                                               # odd coil count: a half pair
    (2, 32, {"CGS_INTERP_CS": "4", "CGS_INTERP_CG": "32"}),                       # ... in 512-thread blocks
    (2, 5, {"CGS_INTERP_CS": "4", "CGS_INTERP_CG": "16"}),
    (2, 12, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0", "CGS_GRID_BATCHED": "0"}),
    (3, 16, {"CGS_GRID_MFMA": "0"}),                                              # 3D register tiles
    (3, 16, {"CGS_INTERP_LDS": "0"}),
    (3, 16, {"CGS_INTERP_CS": "1"}),                                              # lanes = (sample, tap plane)
    (3, 6, {"CGS_INTERP_CS": "1", "CGS_INTERP3_THREADS": "256"}),
    (3, 16, {"CGS_INTERP_CS": "0"}),                                              # lanes = (coil, tap row)
    (3, 5, {"CGS_INTERP_CS": "0"}),                                               # odd coil count: a lone last coil
    (2, 5, {"CGS_INTERP_CS": "0"}),
    (3, 6, {"CGS_INTERP_CS": "2", "CGS_INTERP3_THREADS": "512"}),
    (3, 6, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0"}),
    (2, 32, {"CGS_GRID_BIN": "1"}),                                               # binned LDS scatter
    (2, 32, {"CGS_GRID_BIN": "1", "CGS_BIN_K": "64"}),                            # ... heavy bins split
    (2, 12, {"CGS_GRID_BIN": "1", "CGS_BIN_CG": "4", "CGS_BIN_THREADS": "256"}),  # 4-coil groups, padded
    (2, 32, {"CGS_GRID_BIN": "1", "CGS_BIN_TB": "16", "CGS_BIN_THREADS": "512"}),
    (3, 16, {"CGS_GRID_BIN": "1"}),
    (3, 6, {"CGS_GRID_BIN": "1", "CGS_BIN_K": "256", "CGS_BIN_CG": "2"}),
    (3, 16, {"CGS_GRID_BIN": "1", "CGS_BIN_TB": "8"}),
])
def test_kernel_variants_match_default(cgs, ora, d, C, env):
    """Every alternative gridding / gather kernel that the plan's switches
    select gives E and E^H of the default kernels (radial / koosh trajectories
    with dense centre tiles that are split across work items), and is as close
    to the FP64 oracle as the default.  Kernels that sum in another order differ
    from the default by FP32 rounding, most where thousands of samples meet
    (the koosh-ball centre: ~1e-6)."""
    from codeare_amd import data
    from helpers import rand_img
    rng = np.random.default_rng(7 * d + C)
    if d == 2:
        N = 64
        k = data.radial_2d(90, 2 * N, True)
    else:
        N = 32
        k = data.koosh_3d(400, 2 * N)
    M = k.shape[1]
    w = rng.uniform(0.1, 1.0, M).astype(np.float32)
    S = data.coil_maps(N, d, C, rng)
    x = rand_img(rng, (N,) * d)
    y = rand_img(rng, (M * C,))
    Pv = _plan_env(cgs, env, (N,) * d, M, 2.0 if d == 2 else 1.5, 3, C, k, w, S, None)
    Pd = _plan_env(cgs, {}, (N,) * d, M, 2.0 if d == 2 else 1.5, 3, C, k, w, S, None)
    if env.get("CGS_GRID_BIN") == "1":
        assert Pv.gridding_kind() == "binned"
    assert Pd.gridding_kind() == "tiles"
    xv, xd = Pv.EH(y), Pd.EH(y)
    e1 = rel(xv, xd)
    e2 = rel(Pv.E(x), Pd.E(x))
    O = ora.NFFT((N,) * d, M, 2.0 if d == 2 else 1.5, 3)
    O.kspace(k)
    xo = O.EH_true(S, w, np.asfortranarray(y.reshape((M, C), order="F")))
    ev, ed = rel(xv, xo), rel(xd, xo)
    print(f"d={d} C={C} {env}: E^H {e1:.2e}, E {e2:.2e}; E^H vs oracle {ev:.2e} (default {ed:.2e})", flush=True)
    assert e1 <= 3e-6 and e2 <= 1e-6
    assert ev <= 1e-5 and ev <= 2.0 * ed + 1e-6


def test_time_toeplitz_conv_leaves_the_plan_alone(cgs):
    """cgs_time_toeplitz_conv times the convolution pass on a scratch buffer:
    a positive duration, and the next solve equals the one before it bit for
    bit; without a prepared spectrum it refuses."""
    prob, P = _plan(cgs, 128, 16, 64, 256)   # a shape of the lean conv kernel
    from codeare_amd import data
    y = data.make_data(prob, P.E)
    with pytest.raises(cgs.CgsError):
        P.time_toeplitz_conv(reps=5)
    x1, r1 = P.solve(y, maxit=20, tol=0.0, lam=1e-2, toeplitz=True)
    us = P.time_toeplitz_conv(reps=20)
    x2, r2 = P.solve(y, maxit=20, tol=0.0, lam=1e-2, toeplitz=True)
    print(f"Toeplitz conv at 128^2 x 16 coils: {us:.2f} us per launch")
    assert 0.0 < us < 1e4
    assert np.array_equal(x1, x2) and r1["iters"][0] == r2["iters"][0] == 20
