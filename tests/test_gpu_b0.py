"""GPU parity of the off-resonance (B0) NFFT: codeare's NFFT<cxfl> with "b0"
and "timing" (NFFT.hpp:117-173, NFFT3's mri_inh_3d plan), against the FP64
restatement of the same algorithm (oracle/b0.py; 1e-5) and the direct sum of
the model (the method's own error: 3e-4 at m = 3); NFFT::Adjoint (ITrafo) on
the model against the dense CGNR on the oracle's assembled operator."""
import numpy as np
import pytest

from helpers import rel, rand_traj
from oracle import b0 as B
from oracle import ndft

pytestmark = pytest.mark.gpu


def _case(seed=1, N=32, M=2500, span=10e-3, amp=120.0):
    rng = np.random.default_rng(seed)
    k = rand_traj(rng, 2, M)
    t = np.sort(rng.uniform(0, span, M)).astype(np.float32)
    i0, i1 = np.meshgrid(np.arange(N) - N / 2, np.arange(N) - N / 2, indexing="ij")
    b0 = np.asfortranarray((amp * np.sin(i0 / 7.0) + 0.6 * amp * np.cos(i1 / 4.0)).astype(np.float32))
    x = np.asfortranarray((rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))).astype(np.complex64))
    return rng, k, t, b0, x


@pytest.mark.parametrize("m,C", [(3, 1), (3, 4), (4, 2)])
def test_b0_nfft_forward_adjoint(cgs, ora, m, C):
    rng, k, t, b0, x = _case()
    N, M = x.shape[0], t.size
    P = cgs.Plan((N, N), M, 2.0, m, ncoils=C)
    P.kspace(k)
    P.weights(np.ones(M, np.float32))
    P.b0(b0, t)
    O = B.B0NFFT((N, N), k, t, b0, 2.0, m)
    info = P.b0_info()
    assert info["N3"] == O.N3 and abs(info["omega"] - O.omega) <= 1e-9 * O.omega
    xs = np.asfortranarray(np.stack([x * (1 + 0.5j * c) for c in range(C)], axis=-1)) if C > 1 else x
    f = P.nfft_forward(xs)
    y = (rng.normal(size=(M, C)) + 1j * rng.normal(size=(M, C))).astype(np.complex64)
    a = P.nfft_adjoint(np.asfortranarray(y) if C > 1 else y[:, 0])
    for c in range(C):
        xc = xs[..., c] if C > 1 else xs
        fc = f[:, c] if C > 1 else f
        ac = a[..., c] if C > 1 else a
        e_f, e_a = rel(fc, O.trafo(xc)), rel(ac, O.adjoint(y[:, c]))
        d_f = rel(fc, B.ndft_b0(k, t, b0, xc))
        print(f"m={m} C={C} coil {c}: N3 {O.N3}; forward {e_f:.2e} adjoint {e_a:.2e} vs the FP64 model; "
              f"forward vs direct sum {d_f:.2e}")
        assert e_f <= 1e-5 and e_a <= 1e-5
        assert d_f <= (3e-4 if m == 3 else 2e-5)


def test_b0_nfft_inverse_ref(cgs, ora):
    """NFFT::Adjoint (NFFT.hpp:370-436) on the model: the weighted, damped CGNR
    against the dense twin on the oracle's assembled B0 operator."""
    rng, k, t, b0, x = _case(seed=4, N=12, M=400, span=6e-3, amp=90.0)
    N, M = 12, t.size
    w = rng.uniform(0.2, 1.0, M).astype(np.float32)
    P = cgs.Plan((N, N), M, 2.0, 3, ncoils=1)
    P.kspace(k)
    P.weights(w)
    P.b0(b0, t)
    O = B.B0NFFT((N, N), k, t, b0, 2.0, 3)
    A = np.zeros((M, N * N), dtype=np.complex128)
    for i in range(N * N):
        e = np.zeros(N * N, dtype=np.complex128)
        e[i] = 1.0
        A[:, i] = O.trafo(e.reshape((N, N), order="F"))
    what = ora.damping_mask(2, [N, N]).reshape(-1, order="F")
    y = (rng.normal(size=M) + 1j * rng.normal(size=M)).astype(np.complex64)
    for maxit in (1, 3):
        g = P.nfft_inverse_ref(y, maxit=maxit, eps=0.0)
        o, it = ndft.cgnr_dense(A, y.astype(np.complex128), w.astype(np.float64), what, maxit, 0.0)
        e = rel(g.reshape(-1, order="F"), o)
        print(f"ITrafo on the B0 model, maxit {maxit}: {e:.2e}")
        assert it == maxit and e <= 1e-5


def test_b0_facade_and_limits(cgs):
    """The NFFT facade with b0 / timing params; the NCSENSE operators refuse a
    plan with b0 (codeare's NCSENSE passes b0 to none of its NFFTs); clearing b0
    restores the plain NFFT."""
    from codeare_amd import data
    rng, k, t, b0, x = _case(seed=5)
    N, M = x.shape[0], t.size
    ft = cgs.NFFT({"nk": M, "imsz": (N, N), "m": 3, "alpha": 2.0, "b0": b0, "timing": t})
    ft.KSpace(k)
    ft.Weights(np.ones(M, np.float32))
    O = B.B0NFFT((N, N), k, t, b0, 2.0, 3)
    assert rel(ft * x, O.trafo(x)) <= 1e-5
    P = cgs.Plan((N, N), M, 2.0, 3, ncoils=2)
    P.kspace(k)
    P.weights(np.ones(M, np.float32))
    P.sensitivities(data.coil_maps(N, 2, 2, rng))
    plain = P.nfft_forward(x)
    P.b0(b0, t)
    with pytest.raises(cgs.CgsError) as e:
        P.E(x)
    assert e.value.code == 5
    P.b0(None, None)
    assert rel(P.nfft_forward(x), plain) == 0.0
    P.E(x)
