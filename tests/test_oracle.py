"""Pins the FP64 C oracle (oracle/cgs_oracle.c) without the reference binary.

codeare cannot be built here (NFFT3, FFTW3, Boost and BLAS headers are absent)
and ships no golden vectors for this path, so each oracle function is checked
against what the mathematics fixes: the direct NDFT matrix (independent numpy
code), numpy's FFT, scipy's I0, closed forms of the damping mask and intensity
map, and dense-matrix twins of the CGNR and CG recurrences.
"""
import math

import numpy as np
import pytest
import scipy.special

from helpers import rel, rand_img, rand_traj
from oracle import ndft


def test_bessel_i0(ora):
    for x in [0.0, 1e-3, 0.5, 3.0, 9.42, 14.137, 20.0]:
        assert abs(ora.bessel_i0(x) - scipy.special.i0(x)) <= 1e-13 * scipy.special.i0(x)


def test_kb_window_shape(ora):
    m, b = 3, math.pi * 1.5
    # even, continuous through |t| = m with limit b/pi, sin branch beyond
    for t in [0.3, 1.7, 2.9]:
        assert abs(ora.kb_phi(t, m, b) - ora.kb_phi(-t, m, b)) < 1e-12
    at = ora.kb_phi(3.0, m, b)
    assert abs(at - b / math.pi) < 1e-12
    assert abs(ora.kb_phi(3.0 - 1e-7, m, b) - at) < 1e-5
    assert abs(ora.kb_phi(3.0 + 1e-7, m, b) - at) < 1e-5
    # sinh branch closed form at t = 0
    assert abs(ora.kb_phi(0.0, m, b) - math.sinh(b * m) / (math.pi * m)) < 1e-9 * ora.kb_phi(0.0, m, b)


@pytest.mark.parametrize("n", [2, 3, 4, 8, 12, 48, 64, 96, 192, 256, 384, 512, 60, 7])
def test_fft_vs_numpy(ora, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.abs(ora.fft1d(x, -1) - np.fft.fft(x)).max() <= 1e-12 * n
    assert np.abs(ora.fft1d(x, +1) - np.fft.ifft(x) * n).max() <= 1e-12 * n


def test_ndft_c_vs_numpy(ora):
    rng = np.random.default_rng(1)
    for d, N in [(2, 12), (3, 6)]:
        k = rand_traj(rng, d, 200)
        img = rand_img(rng, (N,) * d).astype(np.complex128)
        A = ndft.ndft_matrix(k, (N,) * d)
        assert rel(ora.ndft_trafo(k, img), A @ img.reshape(-1, order="F")) <= 1e-13
        f = rng.normal(size=200) + 1j * rng.normal(size=200)
        assert rel(ora.ndft_adjoint(k, f, (N,) * d).reshape(-1, order="F"), A.conj().T @ f) <= 1e-13


@pytest.mark.parametrize("d,N,sigma,m,tol", [
    (2, 16, 2.0, 3, 2e-5), (2, 16, 2.0, 6, 1e-10), (2, 16, 1.5, 3, 2e-4),
    (3, 8, 2.0, 3, 3e-5), (3, 8, 2.0, 6, 1e-10), (2, 14, 2.0, 4, 2e-6),
])
def test_nfft_vs_dense_ndft(ora, d, N, sigma, m, tol):
    """The KB NFFT approximates the NDFT with codeare's axis pairing; a wrong
    scaling, sign, axis swap or tap offset fails by O(1)."""
    rng = np.random.default_rng(100 + d + m)
    M = 400
    k = rand_traj(rng, d, M)
    img = rand_img(rng, (N,) * d).astype(np.complex128)
    P = ora.NFFT((N,) * d, M, sigma, m)
    P.kspace(k)
    A = ndft.ndft_matrix(k, (N,) * d)
    assert rel(P.trafo(img), A @ img.reshape(-1, order="F")) <= tol
    f = rng.normal(size=M) + 1j * rng.normal(size=M)
    assert rel(P.adjoint(f).reshape(-1, order="F"), A.conj().T @ f) <= tol


def test_nfft_error_decreases_with_m(ora):
    rng = np.random.default_rng(9)
    k = rand_traj(rng, 2, 300)
    img = rand_img(rng, (16, 16)).astype(np.complex128)
    A = ndft.ndft_matrix(k, (16, 16))
    ref = A @ img.reshape(-1, order="F")
    errs = []
    for m in range(1, 7):
        P = ora.NFFT((16, 16), 300, 2.0, m)
        P.kspace(k)
        errs.append(rel(P.trafo(img), ref))
    assert all(errs[i + 1] < errs[i] for i in range(len(errs) - 1)), errs


def test_nfft_special_cases(ora):
    N = 16
    # one sample at the k-space origin: f = sum of the image
    k0 = np.zeros((2, 1), np.float32)
    img = rand_img(np.random.default_rng(2), (N, N)).astype(np.complex128)
    P = ora.NFFT((N, N), 1, 2.0, 6)
    P.kspace(k0)
    assert abs(P.trafo(img)[0] - img.sum()) <= 1e-9 * abs(img.sum())
    # a delta at image index (i0, i1) gives exp(-2 pi i [(i0-N/2) k(1,j) + (i1-N/2) k(0,j)])
    rng = np.random.default_rng(3)
    k = rand_traj(rng, 2, 50)
    P = ora.NFFT((N, N), 50, 2.0, 6)
    P.kspace(k)
    delta = np.zeros((N, N), np.complex128, order="F")
    i0, i1 = 3, 11
    delta[i0, i1] = 1.0
    expect = np.exp(-2j * np.pi * ((i0 - N // 2) * k[1].astype(np.float64) + (i1 - N // 2) * k[0].astype(np.float64)))
    assert np.abs(P.trafo(delta) - expect).max() <= 1e-9


def test_nfft_adjoint_identity(ora):
    rng = np.random.default_rng(4)
    for d, N in [(2, 24), (3, 8)]:
        M = 500
        k = rand_traj(rng, d, M)
        P = ora.NFFT((N,) * d, M, 2.0, 3)
        P.kspace(k)
        x = rand_img(rng, (N,) * d).astype(np.complex128)
        y = rng.normal(size=M) + 1j * rng.normal(size=M)
        lhs = np.vdot(y, P.trafo(x))
        rhs = np.vdot(P.adjoint(y).reshape(-1, order="F"), x.reshape(-1, order="F"))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_damping_mask_closed_form(ora):
    for N in [16, 32, 64]:
        m = ora.damping_mask(2, [N, N])
        nh = N // 2
        assert m[nh, nh] == 1 and m[0, 0] == 0 and m[0, nh] == 1 and m[nh, 0] == 1
        assert m[1, 1] == 0
        # points at distance exactly nh are inside; just beyond are outside
        assert m[nh + nh - 1, nh] == 1
        count = m.sum()
        assert abs(count - math.pi * nh * nh) <= 4 * N
        assert np.array_equal(m, m.T)
    m3 = ora.damping_mask(3, [8, 8, 8])
    assert m3[4, 4, 4] == 1 and m3[0, 0, 0] == 0 and m3[0, 4, 4] == 1 and m3[0, 0, 4] == 0


def test_intensity_map(ora):
    rng = np.random.default_rng(8)
    S = (rng.normal(size=(8, 8, 5)) + 1j * rng.normal(size=(8, 8, 5))).astype(np.complex64)
    I = ora.intensity_map(np.asfortranarray(S))
    expect = 1.0 / (np.sqrt(np.sum(np.abs(S.astype(np.complex128)) ** 2, axis=-1)) + 1e-9)
    assert np.abs(I - expect).max() <= 1e-6 * expect.max()
    Sn = S / np.sqrt(np.sum(np.abs(S) ** 2, axis=-1, keepdims=True))
    In = ora.intensity_map(np.asfortranarray(Sn.astype(np.complex64)))
    assert np.abs(In - 1.0 / (1.0 + 1e-9)).max() <= 1e-6


def _tiny_cgnr_setup(rng, N=12, M=180):
    k = rand_traj(rng, 2, M)
    w = rng.uniform(0.2, 1.0, M)
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    return k, w, y


def test_itrafo_vs_dense_twin(ora):
    rng = np.random.default_rng(12)
    N, M = 12, 180
    k, w, y = _tiny_cgnr_setup(rng, N, M)
    what = ora.damping_mask(2, [N, N])
    P = ora.NFFT((N, N), M, 2.0, 6)   # m = 6: NFFT == NDFT to 1e-10
    P.kspace(k)
    A = ndft.ndft_matrix(k, (N, N))
    for maxit in (0, 1, 2, 4):
        f, it, _ = P.itrafo(y, w, maxit=maxit, eps=0.0, what=what)
        fd, itd = ndft.cgnr_dense(A, y, w, what.reshape(-1, order="F"), maxit, 0.0)
        assert it == itd == maxit
        if maxit == 0:
            assert np.abs(f).max() == 0
        else:
            assert rel(f.reshape(-1, order="F"), fd) <= 1e-8


def _dense_nfft(P, N):
    """Columns A e_i of the oracle's own NFFT (FP64), image index column-major."""
    A = np.zeros((P.M, N * N), dtype=np.complex128)
    for i in range(N * N):
        e = np.zeros(N * N, dtype=np.complex128)
        e[i] = 1.0
        A[:, i] = P.trafo(e.reshape((N, N), order="F"))
    return A


def _krylov_lsq(B, rhs, k):
    """argmin ||rhs - B u|| over u in K_k(B^H B, B^H rhs) (orthonormal Arnoldi
    basis with re-orthogonalisation, then a dense least-squares solve)."""
    G = B.conj().T @ B
    s = B.conj().T @ rhs
    Q = [s / np.linalg.norm(s)]
    for _ in range(k - 1):
        v = G @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - np.vdot(q, v) * q
        Q.append(v / np.linalg.norm(v))
    Q = np.stack(Q, axis=1)
    c, *_ = np.linalg.lstsq(B @ Q, rhs, rcond=None)
    return Q @ c


def test_itrafo_is_the_krylov_minimiser(ora):
    """NFFT3's weighted damped CGNR (NFFTTraits.hpp:122-138) is CGLS on
    B = W^1/2 A W_hat^1/2 in u, with f_hat = W_hat^1/2 u: its k-th iterate
    minimises ||W^1/2 (y - A f_hat)|| over f_hat in W_hat K_k(B^H B, B^H W^1/2 y).
    A is the oracle's own FP64 NFFT (m = 3, sigma = 2) assembled column by
    column, so this pins the recurrence -- where w and w_hat enter, the step
    and direction updates -- without restating it.  Run long, the iterate is
    the masked weighted least-squares solution."""
    rng = np.random.default_rng(12)
    N, M = 12, 180
    k, w, y = _tiny_cgnr_setup(rng, N, M)
    P = ora.NFFT((N, N), M, 2.0, 3)
    P.kspace(k)
    A = _dense_nfft(P, N)
    f = rng.normal(size=M) + 1j * rng.normal(size=M)
    assert rel(P.adjoint(f).reshape(-1, order="F"), A.conj().T @ f) <= 1e-12   # the CGNR's A^H is A's adjoint
    what = ora.damping_mask(2, [N, N]).reshape(-1, order="F")
    assert 0 < what.sum() < N * N                                                 # corners are damped to zero
    sw = np.sqrt(w)
    B = sw[:, None] * A * np.sqrt(what)[None, :]
    for kk in (1, 2, 3, 4, 6):
        fo, it, _ = P.itrafo(y, w, maxit=kk, eps=0.0)
        assert it == kk
        fk = np.sqrt(what) * _krylov_lsq(B, sw * y, kk)
        assert rel(fo.reshape(-1, order="F"), fk) <= 1e-11
    mask = what > 0
    us, *_ = np.linalg.lstsq(B[:, mask], sw * y, rcond=None)
    fs = np.zeros(N * N, dtype=np.complex128)
    fs[mask] = us
    fo, it, _ = P.itrafo(y, w, maxit=400, eps=0.0)
    assert rel(fo.reshape(-1, order="F"), fs) <= 1e-6


def test_itrafo_krylov_check_catches_a_misplaced_damping(ora):
    """The check above has teeth: the same dense CGNR with w_hat dropped from
    the f_hat update (a plausible slip) leaves the Krylov minimiser at once."""
    rng = np.random.default_rng(12)
    N, M = 12, 180
    k, w, y = _tiny_cgnr_setup(rng, N, M)
    P = ora.NFFT((N, N), M, 2.0, 3)
    P.kspace(k)
    A = _dense_nfft(P, N)
    what = ora.damping_mask(2, [N, N]).reshape(-1, order="F")
    sw = np.sqrt(w)
    B = sw[:, None] * A * np.sqrt(what)[None, :]
    for kk in (2, 3):
        good, _ = ndft.cgnr_dense(A, y, w, what, kk, 0.0)
        # planted mistake: f_hat += alpha p instead of alpha w_hat p
        f = np.zeros(N * N, dtype=np.complex128)
        r = y.copy()
        z = A.conj().T @ (w * r)
        dot_z = np.sum(what * np.abs(z) ** 2)
        p = z.copy()
        for _ in range(kk):
            v = A @ (what * p)
            alpha = dot_z / np.sum(w * np.abs(v) ** 2)
            f = f + alpha * p
            r = r - alpha * v
            z = A.conj().T @ (w * r)
            dz_new = np.sum(what * np.abs(z) ** 2)
            p = (dz_new / dot_z) * p + z
            dot_z = dz_new
        fk = np.sqrt(what) * _krylov_lsq(B, sw * y, kk)
        assert rel(good, fk) <= 1e-11
        assert rel(f, fk) >= 1e-2


def test_itrafo_eps_is_absolute(ora):
    rng = np.random.default_rng(13)
    N, M = 12, 180
    k, w, y = _tiny_cgnr_setup(rng, N, M)
    P = ora.NFFT((N, N), M, 2.0, 3)
    P.kspace(k)
    dot_r0 = float(np.sum(w * np.abs(y) ** 2))
    f, it, hist = P.itrafo(y, w, maxit=5, eps=dot_r0 * 1.0001)
    assert it == 0 and np.abs(f).max() == 0
    f, it, hist = P.itrafo(y, w, maxit=5, eps=dot_r0 * 0.9999)
    assert it >= 1
    # weighted residual decreases monotonically
    f, it, hist = P.itrafo(y, w, maxit=6, eps=0.0)
    assert all(hist[i + 1] <= hist[i] * (1 + 1e-12) for i in range(len(hist) - 1))


def _tiny_sense(rng, N=10, M=160, C=3):
    from codeare_amd import data
    k = rand_traj(rng, 2, M)
    w = rng.uniform(0.2, 1.0, M).astype(np.float32)
    S = data.coil_maps(N, 2, C, rng)
    x = rand_img(rng, (N, N))
    return k, w, S, x


def test_ncsense_E_EH_vs_dense(ora):
    rng = np.random.default_rng(21)
    N, M, C = 10, 160, 3
    k, w, S, x = _tiny_sense(rng, N, M, C)
    P = ora.NFFT((N, N), M, 2.0, 6)
    P.kspace(k)
    Ed = ndft.dense_E(k, (N, N), S.astype(np.complex128))
    y = P.E(S, x)
    assert rel(y.reshape(-1, order="F"), Ed @ x.reshape(-1, order="F").astype(np.complex128)) <= 1e-6
    yy = rand_img(rng, (M, C))
    wC = np.tile(w.astype(np.float64), C)
    xh = P.EH_true(S, w, yy)
    assert rel(xh.reshape(-1, order="F"), Ed.conj().T @ (wC * yy.reshape(-1, order="F"))) <= 1e-6


def test_ncsense_EH_ref_vs_dense(ora):
    """reference E^H = I .* sum_c conj(S_c) .* CGNR_c(y_c) with the dense CGNR."""
    rng = np.random.default_rng(22)
    N, M, C = 10, 160, 3
    k, w, S, x = _tiny_sense(rng, N, M, C)
    P = ora.NFFT((N, N), M, 2.0, 6)
    P.kspace(k)
    yy = rand_img(rng, (M, C))
    xr = P.EH_ref(S, w, yy, ftiter=2, fteps=0.0)
    A = ndft.ndft_matrix(k, (N, N))
    what = ora.damping_mask(2, [N, N]).reshape(-1, order="F")
    acc = np.zeros(N * N, np.complex128)
    Sf = S.reshape(-1, C, order="F").astype(np.complex128)
    for c in range(C):
        fc, _ = ndft.cgnr_dense(A, yy[:, c].astype(np.complex128), w.astype(np.float64), what, 2, 0.0)
        acc += np.conj(Sf[:, c]) * fc
    I = 1.0 / (np.sqrt(np.sum(np.abs(Sf) ** 2, axis=1)) + 1e-9)
    assert rel(xr.reshape(-1, order="F"), acc * I) <= 1e-5


def test_cg_true_vs_dense(ora):
    rng = np.random.default_rng(31)
    N, M, C = 10, 200, 3
    k, w, S, x = _tiny_sense(rng, N, M, C)
    P = ora.NFFT((N, N), M, 2.0, 6)
    P.kspace(k)
    y = P.E(S, x)
    Ed = ndft.dense_E(k, (N, N), S.astype(np.complex128))
    W = np.tile(w.astype(np.float64), C)
    AHA = Ed.conj().T @ (W[:, None] * Ed)
    b = Ed.conj().T @ (W * y.reshape(-1, order="F").astype(np.complex128))
    lam = 1e-3
    xo, it, hist = P.cg_true(S, w, y, maxit=8, tol=0.0, lam=lam)
    xd, itd, histd = ndft.cg_normal_dense(AHA, b, lam, 8, 0.0)
    assert it == itd == 8
    assert rel(xo.reshape(-1, order="F"), xd) <= 1e-4
    # run to convergence: the normal-equation solution
    xo, it, hist = P.cg_true(S, w, y, maxit=200, tol=1e-6, lam=lam)
    xs = np.linalg.solve(AHA + lam * np.eye(N * N), b)
    assert hist[-1] <= 1e-6
    assert rel(xo.reshape(-1, order="F"), xs) <= 1e-4
    # warm start from the solution stops at once
    xw, itw, _ = P.cg_true(S, w, y, maxit=50, tol=1e-3, lam=lam, x0=xs.reshape((N, N), order="F"))
    assert itw == 0


def test_cgls_ref_semantics(ora):
    """CGLS.hpp:30-59: maxit 0 returns A/y; squared-ratio stop."""
    rng = np.random.default_rng(41)
    N, M, C = 10, 200, 3
    k, w, S, x = _tiny_sense(rng, N, M, C)
    P = ora.NFFT((N, N), M, 2.0, 3)
    P.kspace(k)
    y = P.E(S, x)
    x0, it, _ = P.cgls_ref(S, w, y, mode=1, maxit=0)
    assert it == 0
    assert rel(x0, P.EH_ref(S, w, y)) == 0.0
    # true-adjoint CGLS equals the FP64-scalar CG at equal iteration counts
    xa, ita, ha = P.cgls_ref(S, w, y, mode=0, maxit=6, cgeps=0.0, lam=1e-3)
    xb, itb, hb = P.cg_true(S, w, y, maxit=6, tol=0.0, lam=1e-3)
    assert ita == itb == 6
    assert rel(xa, xb) <= 1e-4
    assert np.allclose(ha[: len(hb)], hb[: len(ha)] ** 2, rtol=1e-3, atol=1e-9)
