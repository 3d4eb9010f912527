"""cgs_solve_noalloc (the allocation-free solve) against cgs_solve on the same
plans: equal iteration counts, residual histories and images, on two calls in
a row of which the second is the smaller problem (the shared scratch keeps the
first call's contents), and the limits it reports instead of allocating."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _problem(cgs, N, C, spokes, F=1, seed=0):
    import torch
    from codeare_amd import data
    ks, counts, first = [], [], 0
    for f in range(F):
        c = spokes + 3 * f
        ks.append(data.radial_2d(c, 2 * N, True, first))
        counts.append(c)
        first += c
    k = np.asfortranarray(np.concatenate(ks, axis=1))
    off = np.concatenate([[0], np.cumsum(np.array(counts) * 2 * N)]).astype(np.int64)
    rng = np.random.default_rng(seed)
    w = data.density(k, "k", 2 * N)
    S = data.coil_maps(N, 2, C, rng)
    P = cgs.Plan((N, N), k.shape[1], 2.0, 3, ncoils=C, frame_offsets=off if F > 1 else None, device_io=True)
    dev = torch.device("cuda:0")
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    x = torch.from_numpy(np.ascontiguousarray(
        (rng.normal(size=N * N * F) + 1j * rng.normal(size=N * N * F)).astype(np.complex64))).to(dev)
    y = torch.empty(k.shape[1] * C, dtype=torch.complex64, device=dev)
    P.E(x, out=y)
    P.sync()
    y += 0.05 * torch.randn_like(y).abs().mean() * torch.randn_like(y)
    torch.cuda.synchronize()
    return P, y


def _both(P, y, **kw):
    import torch
    n = P.Ntot * P.F
    xa = torch.empty(n, dtype=torch.complex64, device=y.device)
    xb = torch.empty_like(xa)
    hist = kw["maxit"] > 0        # zero iterations run no CG (CGLS.hpp:32-33): no history
    _, ra = P.solve(y, out=xa, **kw)
    ha = P.residual_history() if hist else None
    _, rb = P.solve(y, out=xb, noalloc=True, **kw)
    hb = P.residual_history() if hist else None
    torch.cuda.synchronize()
    return xa.cpu().numpy(), ra, ha, xb.cpu().numpy(), rb, hb


CASES = [
    ("toeplitz", dict(maxit=60, tol=1e-5, lam=1e-3, toeplitz=True)),
    ("nufft", dict(maxit=40, tol=1e-5, lam=1e-3)),
    ("reference", dict(maxit=6, tol=1e-10, tol_kind=1, lam=5e-2, eh_mode=1, ftiter=3, fteps=1e-3)),
    ("zero-iterations", dict(maxit=0, tol=0.0, lam=0.0)),
]


@pytest.mark.parametrize("name,kw", CASES)
def test_noalloc_equals_solve_on_two_calls(cgs, name, kw):
    # call 1: 128^2, 16 coils, three frames; call 2: 64^2, 8 coils, one frame
    P1, y1 = _problem(cgs, 128, 16, 40, F=3, seed=1)
    P2, y2 = _problem(cgs, 64, 8, 24, F=1, seed=2)
    if kw.get("toeplitz"):
        import torch
        for P, y in ((P1, y1), (P2, y2)):   # the spectrum is prepared (allocated) outside the allocation-free call
            x = torch.empty(P.Ntot * P.F, dtype=torch.complex64, device=y.device)
            P.solve(y, out=x, maxit=1, tol=0.0, lam=1e-3, toeplitz=True)
    for tag, P, y in (("first (larger)", P1, y1), ("second (smaller)", P2, y2)):
        xa, ra, ha, xb, rb, hb = _both(P, y, **kw)
        e = rel(xb, xa)
        print(f"{name} {tag}: iterations {ra['iters'][:P.F]} / {rb['iters'][:P.F]}, image {e:.2e}")
        assert ra["iters"][:P.F] == rb["iters"][:P.F]
        assert e <= 1e-6
        if kw["maxit"] > 0:
            assert len(ha) == len(hb) and np.allclose(ha, hb, rtol=1e-9, atol=0)
            assert np.allclose(ra["final_rel"][:P.F], rb["final_rel"][:P.F], rtol=1e-9)


def test_noalloc_limits(cgs):
    import torch
    P, y = _problem(cgs, 64, 8, 24)
    x = torch.empty(P.Ntot, dtype=torch.complex64, device=y.device)
    with pytest.raises(cgs.CgsError) as e:          # the spectrum is not prepared yet
        P.solve(y, out=x, maxit=10, tol=0.0, toeplitz=True, noalloc=True)
    assert e.value.code == 2
    with pytest.raises(cgs.CgsError) as e:          # graph replay would instantiate (allocate)
        P.solve(y, out=x, maxit=10, tol=0.0, graph_iters=8, noalloc=True)
    assert e.value.code == 1
    with pytest.raises(cgs.CgsError) as e:          # history beyond the scratch
        P.solve(y, out=x, maxit=(8 << 20) // 8, tol=0.0, noalloc=True)
    assert e.value.code == 1 and "scratch" in str(e.value)
    Ph = cgs.Plan((64, 64), 100, 2.0, 3, ncoils=2)  # host-buffer plan
    with pytest.raises(cgs.CgsError):
        Ph.solve(np.zeros((100, 2), np.complex64), maxit=1, noalloc=True)
