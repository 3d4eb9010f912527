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
