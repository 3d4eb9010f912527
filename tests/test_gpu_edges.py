"""Edge cases of the GPU path against the FP64 oracle: trajectory values on the
boundary of [-1/2, 1/2) and on grid-cell positions, a single sample and coil,
invalid trajectories, and a NaN in one frame's data (codeare's CGLS stops on a
NaN residual, CGLS.hpp:41; here the frame stops, the others go on)."""
import numpy as np
import pytest

from helpers import rel, rand_img

pytestmark = pytest.mark.gpu


def _ops(cgs, ora, k, N, C, sigma, m, w, S):
    M = k.shape[1]
    P = cgs.Plan((N, N), M, sigma, m, ncoils=C)
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    O = ora.NFFT((N, N), M, sigma, m)
    O.kspace(k)
    return P, O


def test_trajectory_boundaries(cgs, ora):
    """Samples at k = -1/2 exactly, at the largest float below +1/2, at 0, and
    exactly on oversampled-grid cell positions (n k integral, where the window's
    taps end on a cell edge): E and E^H W within 1e-5 of the oracle."""
    from codeare_amd import data
    rng = np.random.default_rng(17)
    N, C, sigma, m = 32, 4, 2.0, 3
    n = int(sigma * N)
    hi = np.nextafter(np.float32(0.5), np.float32(0.0))
    special = [(-0.5, -0.5), (-0.5, hi), (hi, -0.5), (hi, hi), (0.0, 0.0), (-0.5, 0.0), (0.0, hi)]
    cells = [(i / n, -j / n) for i, j in [(1, 3), (5, 0), (17, 23), (31, 31), (0, 32)]]
    rand = rng.uniform(-0.5, 0.5, size=(300, 2)).tolist()
    k = np.asfortranarray(np.array(special + cells + rand, dtype=np.float32).T)
    M = k.shape[1]
    w = rng.uniform(0.2, 1.0, M).astype(np.float32)
    S = data.coil_maps(N, 2, C, rng)
    x = rand_img(rng, (N, N))
    y = rand_img(rng, (M * C,))
    P, O = _ops(cgs, ora, k, N, C, sigma, m, w, S)
    e1 = rel(P.E(x), O.E(S, x))
    e2 = rel(P.EH(y), O.EH_true(S, w, np.asfortranarray(y.reshape((M, C), order="F"))))
    print(f"boundary and cell-edge samples: E {e1:.2e}, E^H {e2:.2e}")
    assert e1 <= 1e-5 and e2 <= 1e-5


def test_single_sample_single_coil(cgs, ora):
    from codeare_amd import data
    rng = np.random.default_rng(18)
    N = 16
    for kv in ((0.0, 0.0), (0.23, -0.41), (-0.5, -0.5)):
        k = np.asfortranarray(np.array([kv], dtype=np.float32).T)
        w = np.ones(1, np.float32)
        S = data.coil_maps(N, 2, 1, rng)
        x = rand_img(rng, (N, N))
        y = rand_img(rng, (1,))
        P, O = _ops(cgs, ora, k, N, 1, 2.0, 3, w, S)
        e1 = rel(P.E(x), O.E(S, x))
        e2 = rel(P.EH(y), O.EH_true(S, w, np.asfortranarray(y.reshape((1, 1)))))
        print(f"one sample at {kv}: E {e1:.2e}, E^H {e2:.2e}")
        assert e1 <= 1e-5 and e2 <= 1e-5


@pytest.mark.parametrize("bad", [0.5, -0.5000001, np.nan, np.inf])
def test_invalid_trajectory_is_refused(cgs, bad):
    k = np.zeros((2, 8), dtype=np.float32)
    k[1, 3] = bad
    P = cgs.Plan((16, 16), 8, 2.0, 3, ncoils=1)
    with pytest.raises(cgs.CgsError) as e:
        P.kspace(np.asfortranarray(k))
    assert e.value.code == cgs.CGS_ERR_ARG


@pytest.mark.parametrize("toeplitz", [False, True])
def test_nan_in_one_frame(cgs, toeplitz):
    """Two frames, a NaN in frame 1's data: frame 1 stops at once with the NaN
    status (CGS_ERR_NAN, or an exception with check_nan), frame 0's solve is
    the one it has alone."""
    from codeare_amd import data
    rng = np.random.default_rng(19)
    N, C, R, sp = 32, 4, 64, 24
    k0 = data.radial_2d(sp, R, True, 0)
    k1 = data.radial_2d(sp, R, True, sp)
    k = np.asfortranarray(np.concatenate([k0, k1], axis=1))
    Mf = k0.shape[1]
    off = np.array([0, Mf, 2 * Mf], dtype=np.int64)
    w = data.density(k, "k", R)
    S = data.coil_maps(N, 2, C, rng)
    P = cgs.Plan((N, N), 2 * Mf, 2.0, 3, ncoils=C, frame_offsets=off)
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    x = rand_img(rng, (N, N, 2))
    y = np.asarray(P.E(x)).copy()
    # frame f's block of y is [C*off[f], C*off[f+1]) (samples of the frame, coil-major)
    y[C * Mf + 5] = np.complex64(np.nan)
    opts = dict(maxit=100, tol=1e-5, lam=0.1, toeplitz=toeplitz)
    with pytest.raises(cgs.CgsError) as e:
        P.solve(y, **opts)
    assert e.value.code == cgs.CGS_ERR_NAN
    xs, rep = P.solve(y, check_nan=False, **opts)
    P0 = cgs.Plan((N, N), Mf, 2.0, 3, ncoils=C)
    P0.kspace(np.asfortranarray(k0))
    P0.weights(w[:Mf])
    P0.sensitivities(S)
    x0, r0 = P0.solve(np.asfortranarray(y[:C * Mf].reshape((Mf, C), order="F")), **opts)
    e0 = rel(xs[..., 0], x0)
    print(f"toeplitz={toeplitz}: frame 0 iters {rep['iters'][0]} (alone {r0['iters'][0]}), image {e0:.2e}; "
          f"frame 1 iters {rep['iters'][1]}")
    assert abs(rep["iters"][0] - r0["iters"][0]) <= 1
    assert e0 <= (1e-5 if rep["iters"][0] == r0["iters"][0] else 1e-3)
    assert rep["iters"][1] <= 1
    assert np.all(np.isfinite(xs[..., 0]))
