"""The oracle's joint solvers for frames and Cartesian slices (codeare's CGLS on
the whole (N^d, F) image of nmany frames / 3rd_dim_cart slices), pinned by
what the mathematics fixes:
  * one frame is the single-plan solver, bit for bit;
  * F identical frames give the single-frame iterates on every frame (the
    joint alpha and beta are ratios of F-fold sums);
  * for the linear true-adjoint system, CG's k-th iterate from zero is the
    unique minimiser of the A-norm error over the Krylov space K_k(A, b), here
    computed directly from the assembled block-diagonal matrix -- and it is not
    what k steps of separate per-frame CGs give;
  * with every per-frame map set equal to the shared maps the per-frame-maps
    path reproduces the shared one, and maxit = 0 returns each slice's own
    reference E^H with its own maps.
CPU only."""
import numpy as np
import pytest

from helpers import rel, rand_img, rand_traj


def _setup(ora, counts=(40, 70), N=8, C=2, seed=1, same=False, per_frame_maps=False):
    rng = np.random.default_rng(seed)
    F = len(counts)
    k0 = rand_traj(rng, 2, counts[0])
    plans, ks = [], []
    for M in counts:
        k = k0 if same else rand_traj(rng, 2, M)
        O = ora.NFFT((N, N), M, 2.0, 3)
        O.kspace(k)
        plans.append(O)
        ks.append(k)
    w0 = rng.uniform(0.2, 1.0, counts[0]).astype(np.float32)
    w = np.concatenate([w0 if same else rng.uniform(0.2, 1.0, M).astype(np.float32) for M in counts])
    S0 = rand_img(rng, (N, N, C))
    if per_frame_maps:
        S = np.stack([S0 if same else rand_img(rng, (N, N, C)) for _ in range(F)], axis=2)
    else:
        S = S0
    y0 = rand_img(rng, (counts[0], C))
    ys = [y0 if same else rand_img(rng, (M, C)) for M in counts]
    y = np.concatenate([b.reshape(-1, order="F") for b in ys])
    moff = np.concatenate([[0], np.cumsum(counts)])
    return plans, S, w, ys, y, moff


def _maps(S, f, per_frame):
    return np.asfortranarray(S[:, :, f, :]) if per_frame else S


def test_one_frame_is_the_single_plan_solver(ora):
    plans, S, w, ys, y, _ = _setup(ora, counts=(50,))
    xa, ia, ha = ora.cgls_ref_frames(plans, S, w, y, mode=1, maxit=4, cgeps=1e-9, lam=0.05)
    xb, ib, hb = plans[0].cgls_ref(S, w, ys[0], mode=1, maxit=4, cgeps=1e-9, lam=0.05)
    assert ia == ib and np.array_equal(xa[..., 0], xb) and np.array_equal(ha, hb)
    xa, ia, ha = ora.cg_true_frames(plans, S, w, y, maxit=6, tol=0.0, lam=0.1)
    xb, ib, hb = plans[0].cg_true(S, w, ys[0], maxit=6, tol=0.0, lam=0.1)
    assert ia == ib and np.array_equal(xa[..., 0], xb) and np.array_equal(ha, hb)


@pytest.mark.parametrize("per_frame_maps", [False, True])
def test_identical_frames_replicate_the_single_frame(ora, per_frame_maps):
    plans, S, w, ys, y, _ = _setup(ora, counts=(50, 50, 50), same=True, per_frame_maps=per_frame_maps)
    S1 = _maps(S, 0, per_frame_maps)
    xa, ia, ha = ora.cgls_ref_frames(plans, S, w, y, mode=1, maxit=4, cgeps=1e-9, lam=0.05,
                                     maps_per_frame=per_frame_maps)
    xb, ib, hb = plans[0].cgls_ref(S1, w[:50], ys[0], mode=1, maxit=4, cgeps=1e-9, lam=0.05)
    assert ia == ib
    for f in range(3):
        assert rel(xa[..., f], xb) <= 1e-5
    xa, ia, ha = ora.cg_true_frames(plans, S, w, y, maxit=8, tol=0.0, lam=0.1, maps_per_frame=per_frame_maps)
    xb, ib, hb = plans[0].cg_true(S1, w[:50], ys[0], maxit=8, tol=0.0, lam=0.1)
    assert ia == ib and np.allclose(ha, hb, rtol=1e-5)
    for f in range(3):
        assert rel(xa[..., f], xb) <= 1e-5


@pytest.mark.parametrize("per_frame_maps", [False, True])
def test_joint_cg_is_the_krylov_minimiser(ora, per_frame_maps):
    plans, S, w, ys, y, moff = _setup(ora, counts=(40, 70), per_frame_maps=per_frame_maps, seed=4)
    N, lam, K = 8, 0.3, 4
    Nt = N * N
    F = len(plans)
    # assemble A = blockdiag(E_f^H W_f E_f) + lam I and b = E^H W y column by column
    A = np.zeros((F * Nt, F * Nt), dtype=np.complex128)
    b = np.zeros(F * Nt, dtype=np.complex128)
    for f, O in enumerate(plans):
        Sf = _maps(S, f, per_frame_maps)
        wf = w[moff[f]:moff[f + 1]]
        for i in range(Nt):
            e = np.zeros((N, N), dtype=np.complex64, order="F")
            e.reshape(-1, order="F")[i] = 1.0
            col = O.EH_true(Sf, wf, O.E(Sf, e)).reshape(-1, order="F")
            A[f * Nt:(f + 1) * Nt, f * Nt + i] = col
        b[f * Nt:(f + 1) * Nt] = O.EH_true(Sf, wf, ys[f]).reshape(-1, order="F")
    A += lam * np.eye(F * Nt)
    A = 0.5 * (A + A.conj().T)   # Hermitian up to the operators' FP32 rounding
    kry = [b]
    for _ in range(K - 1):
        kry.append(A @ kry[-1])
    V, _ = np.linalg.qr(np.stack(kry, axis=1))
    xk = V @ np.linalg.solve(V.conj().T @ A @ V, V.conj().T @ b)
    xk = xk.reshape((F, Nt)).T.reshape((N, N, F), order="F")
    # the joint solvers (true adjoint): FP64-scalar CG and codeare's CGLS in mode 0
    xj, it, _ = ora.cg_true_frames(plans, S, w, y, maxit=K, tol=0.0, lam=lam, maps_per_frame=per_frame_maps)
    assert it == K
    ej = rel(xj, xk)
    xc, itc, _ = ora.cgls_ref_frames(plans, S, w, y, mode=0, maxit=K, cgeps=0.0, lam=lam,
                                     maps_per_frame=per_frame_maps)
    ec = rel(xc, xk)
    # separate per-frame CGs reach a different point of the (larger) sum of the
    # per-frame Krylov spaces
    xs = np.stack([plans[f].cg_true(_maps(S, f, per_frame_maps), w[moff[f]:moff[f + 1]], ys[f], maxit=K,
                                    tol=0.0, lam=lam)[0] for f in range(F)], axis=2)
    es = rel(xs, xk)
    print(f"joint CG vs Krylov minimiser {ej:.2e}, joint CGLS {ec:.2e}, per-frame CGs {es:.2e}")
    assert ej <= 1e-4 and ec <= 1e-4
    assert es >= 100 * max(ej, ec)


def test_per_frame_maps_layout(ora):
    plans, S, w, ys, y, moff = _setup(ora, counts=(40, 70, 30), per_frame_maps=True, seed=7)
    F = len(plans)
    # maxit = 0: CGLS returns E^H y, slice by slice with each slice's maps
    x0, it, _ = ora.cgls_ref_frames(plans, S, w, y, mode=1, maxit=0, maps_per_frame=True)
    assert it == 0
    for f in range(F):
        xf = plans[f].EH_ref(_maps(S, f, True), w[moff[f]:moff[f + 1]], ys[f])
        assert np.array_equal(x0[..., f], xf)
    # per-frame maps all equal to the shared maps reproduce the shared path
    Ssh = _maps(S, 0, True)
    Srep = np.stack([Ssh] * F, axis=2)
    xa, _, _ = ora.cgls_ref_frames(plans, Srep, w, y, mode=1, maxit=3, cgeps=0.0, lam=0.05, maps_per_frame=True)
    xb, _, _ = ora.cgls_ref_frames(plans, Ssh, w, y, mode=1, maxit=3, cgeps=0.0, lam=0.05)
    assert np.array_equal(xa, xb)
