"""The folded CG iteration (CGS_CG_TAILFOLD=1: the CG tail of codeare's loop,
CGLS.hpp:50-56, moved into the next iteration's pad pass, with ||r_new||^2 from
rn - 2 alpha <r, q> + alpha^2 |q|^2) against the default fused iteration, which
the oracle tests pin: iteration counts within one (or rounding's spread on an
ill-conditioned system), residual histories that agree until rounding parts
them, images that meet the tolerance on the true residual, in launch and graph
mode, with frames that stop at different iterations and solves that end at
maxit."""
import os

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _plan(cgs, tailfold, N, C, k, w, S, foff):
    old = os.environ.get("CGS_CG_TAILFOLD")
    os.environ["CGS_CG_TAILFOLD"] = "1" if tailfold else "0"
    try:
        M = k.shape[1]
        P = cgs.Plan((N, N), M, 2.0, 3, ncoils=C, frame_offsets=foff)
    finally:
        if old is None:
            os.environ.pop("CGS_CG_TAILFOLD", None)
        else:
            os.environ["CGS_CG_TAILFOLD"] = old
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    return P


def _problem(N, C, spokes, frames, seed):
    from codeare_amd import data
    rng = np.random.default_rng(seed)
    ks = [data.radial_2d(spokes, 2 * N, True, f * spokes) for f in range(frames)]
    k = np.asfortranarray(np.concatenate(ks, axis=1))
    Mf = ks[0].shape[1]
    foff = np.arange(frames + 1, dtype=np.int64) * Mf if frames > 1 else None
    w = data.density(k, "k", 2 * N)
    S = data.coil_maps(N, 2, C, rng)
    return k, w, S, foff, rng


def _true_residual(P, y, x, lam):
    """||E^H W y - (E^H W E + lambda) x|| / ||E^H W y||, in FP64 from the
    operator's FP32 outputs (the solver's own residual is a recurrence)."""
    b = np.asarray(P.EH(y)).astype(np.complex128)
    q = np.asarray(P.normal(x, lam, toeplitz=True)).astype(np.complex128)
    if b.ndim == 3:   # per frame
        return [float(np.linalg.norm(b[..., f] - q[..., f]) / np.linalg.norm(b[..., f])) for f in range(b.shape[2])]
    return float(np.linalg.norm(b - q) / np.linalg.norm(b))


@pytest.mark.parametrize("N,C,frames", [(128, 32, 2), (256, 32, 1), (192, 48, 1), (128, 32, 3)])
@pytest.mark.parametrize("graph", [0, 8])
def test_tailfold_matches_fused_iteration(cgs, N, C, frames, graph):
    from helpers import rand_img
    k, w, S, foff, rng = _problem(N, C, 48 if N < 256 else 96, frames, 11 + N + C)
    Pf = _plan(cgs, True, N, C, k, w, S, foff)
    Pd = _plan(cgs, False, N, C, k, w, S, foff)
    x = rand_img(rng, (N, N, frames) if frames > 1 else (N, N))
    y = Pd.E(x)
    # noise in frame 0's data: it stops later than frame 1
    if frames > 1:
        y = np.asarray(y).copy()
        h = y.size // 2
        y[:h] += 0.3 * np.sqrt(np.mean(np.abs(y[:h]) ** 2)) * rand_img(rng, (h,))
    for maxit, tol, lam in ((300, 1e-3, 0.05), (7, 0.0, 0.01), (9, 0.0, 0.0)):
        xf, rf = Pf.solve(y, maxit=maxit, tol=tol, lam=lam, toeplitz=True, graph_iters=graph)
        xd, rd = Pd.solve(y, maxit=maxit, tol=tol, lam=lam, toeplitz=True, graph_iters=graph)
        e = rel(xf, xd)
        print(f"N={N} C={C} F={frames} graph={graph} maxit={maxit} tol={tol}: iterations "
              f"{rf['iters']} / fused {rd['iters']}, image {e:.2e}, final rel {rf['final_rel']} / "
              f"{rd['final_rel']}", flush=True)
        for f in range(frames):
            hf = Pf.residual_history(f)
            hd = Pd.residual_history(f)
            if tol == 0.0:
                assert rf["iters"][f] == rd["iters"][f] == maxit
                # a few iterations: the two orders of rounding have not parted
                assert np.allclose(hf, hd, rtol=1e-4, atol=0), (hf, hd)
            else:
                assert abs(rf["iters"][f] - rd["iters"][f]) <= 1
                n = min(len(hf), len(hd))
                assert n >= min(rf["iters"][f], rd["iters"][f]) + 1
                assert np.all(np.isfinite(hf[:n]))
                # CG on an ill-conditioned system: rounding differences grow
                # with the iterations, so the histories agree closely at first
                assert np.allclose(hf[:30], hd[:30], rtol=1e-2, atol=0), (hf[:30], hd[:30])
                if rf["iters"][f] < maxit:   # stopped at the tolerance
                    assert rf["final_rel"][f] <= tol and rd["final_rel"][f] <= tol
                else:                        # both ran out of iterations: the same residual
                    assert abs(rf["final_rel"][f] / rd["final_rel"][f] - 1.0) <= 0.05
        if tol > 0.0:
            tf, td = _true_residual(Pd, y, xf, lam), _true_residual(Pd, y, xd, lam)
            tf, td = np.atleast_1d(tf), np.atleast_1d(td)
            print(f"  true residual {tf} / fused {td}", flush=True)
            for f in range(frames):
                if rf["iters"][f] < maxit:
                    assert tf[f] <= 2 * tol and td[f] <= 2 * tol
                else:
                    assert abs(tf[f] / td[f] - 1.0) <= 0.05
        # stopping one iteration apart at tol 1e-3 moves the image by about one
        # CG step; the true residual above is the check that matters there
        same = rf["iters"] == rd["iters"]
        assert e <= (1e-4 if tol == 0.0 else 2e-3 if same else 1e-2)
        assert rf["status"] == rd["status"] == 0


def test_tailfold_warm_start_and_repeat(cgs):
    """A warm start runs folded too, a second solve on the same plan (the graph
    replayed) gives the first's result bit for bit, and the image the folded
    solve returns meets the tolerance on the true residual as the fused
    iteration's does, in as many iterations to within rounding's spread."""
    from helpers import rand_img
    N, C = 128, 32
    k, w, S, foff, rng = _problem(N, C, 48, 1, 5)
    Pf = _plan(cgs, True, N, C, k, w, S, foff)
    Pd = _plan(cgs, False, N, C, k, w, S, foff)
    y = Pd.E(rand_img(rng, (N, N)))
    x0 = rand_img(rng, (N, N)) * 0.1
    lam = 0.02
    opts = dict(maxit=400, tol=1e-4, lam=lam, toeplitz=True, graph_iters=4)
    xa, ra = Pf.solve(y, x0=x0, **opts)
    xb, rb = Pf.solve(y, x0=x0, **opts)
    xd, rd = Pd.solve(y, x0=x0, **opts)
    ta, td = _true_residual(Pd, y, xa, lam), _true_residual(Pd, y, xd, lam)
    print(f"warm start: iterations {ra['iters'][0]} / {rb['iters'][0]} / fused {rd['iters'][0]}, "
          f"vs fused {rel(xa, xd):.2e}; true residual {ta:.2e} / fused {td:.2e}", flush=True)
    assert np.array_equal(xa, xb) and ra["iters"] == rb["iters"]
    assert abs(ra["iters"][0] - rd["iters"][0]) <= max(1, int(0.03 * rd["iters"][0]))
    assert ta <= 2e-4 and td <= 2e-4
    assert rel(xa, xd) <= 5e-3
