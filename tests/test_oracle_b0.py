"""The off-resonance (B0) oracle (oracle/b0.py), pinned by direct summation of
f_j = sum_r x(r) exp(-2 pi i k_j.r) exp(-2 pi i b0(r) t_j) and by special cases:
the adjoint identity, a zero field (the plain NDFT), a uniform field (a phase
ramp in time), and the size rule of the third axis."""
import numpy as np
import pytest

from helpers import rel, rand_traj
from oracle import b0 as B
from oracle import ndft


def _case(seed=1, N=16, M=600, span=8e-3, amp=150.0):
    rng = np.random.default_rng(seed)
    k = rand_traj(rng, 2, M)
    t = np.sort(rng.uniform(0, span, M)).astype(np.float32)
    i0, i1 = np.meshgrid(np.arange(N) - N / 2, np.arange(N) - N / 2, indexing="ij")
    b0 = np.asfortranarray((amp * np.sin(i0 / 5.0) + 0.5 * amp * np.cos(i1 / 3.0)).astype(np.float32))
    x = np.asfortranarray(rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N)))
    return rng, k, t, b0, x


@pytest.mark.parametrize("m,tol", [(3, 3e-4), (4, 1e-5)])
def test_b0_model_matches_direct_sum(ora, m, tol):
    rng, k, t, b0, x = _case()
    O = B.B0NFFT((16, 16), k, t, b0, 2.0, m)
    y = rng.normal(size=t.size) + 1j * rng.normal(size=t.size)
    e_f = rel(O.trafo(x), B.ndft_b0(k, t, b0, x))
    e_a = rel(O.adjoint(y), B.ndft_b0_adjoint(k, t, b0, y, (16, 16)))
    assert e_f <= tol and e_a <= tol
    # the fast operator's adjoint is its exact adjoint
    lhs = np.vdot(O.trafo(x), y)
    rhs = np.vdot(x.reshape(-1, order="F"), O.adjoint(y).reshape(-1, order="F"))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_b0_direct_sum_special_cases(ora):
    rng, k, t, b0, x = _case(seed=2)
    A = ndft.ndft_matrix(k, (16, 16))
    # zero field: the plain NDFT
    z = np.zeros_like(b0)
    assert rel(B.ndft_b0(k, t, z, x), A @ x.reshape(-1, order="F")) <= 1e-13
    # uniform field c: the NDFT times exp(-2 pi i c t_j)
    c = np.full_like(b0, 37.5)
    ref = np.exp(-2j * np.pi * 37.5 * t.astype(np.float64)) * (A @ x.reshape(-1, order="F"))
    assert rel(B.ndft_b0(k, t, c, x), ref) <= 1e-12
    # and the fast model reproduces both
    for field in (z, c):
        O = B.B0NFFT((16, 16), k, t, field, 2.0, 3)
        assert rel(O.trafo(x), B.ndft_b0(k, t, field, x)) <= 5e-4   # the window pair's own error at sigma3 = 1.2, m = 3


@pytest.mark.parametrize("span,amp,m", [(8e-3, 150.0, 3), (30e-3, 400.0, 3), (2e-3, 50.0, 2), (0.0, 100.0, 3)])
def test_b0_third_axis_rule(span, amp, m):
    rng, k, t, b0, x = _case(seed=3, span=span, amp=amp)
    N3, omega, ts = B.third_axis(b0, t, m, 2.0)
    assert N3 % 2 == 0 and N3 >= 2 * m + 2
    tau = (t.astype(np.float64) - ts) * omega / N3
    nu = N3 * b0.astype(np.float64) / omega
    assert np.abs(tau).max() <= 0.5 / B.SIGMA3 + 1e-12        # inside the window's pass band
    assert np.all(np.floor(nu - m) >= -N3 // 2) and np.all(np.ceil(nu + m) <= N3 // 2)   # every tap on the axis
