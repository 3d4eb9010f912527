"""GPU parity of codeare's 3rd_dim_cart stack-of-stars path and of the joint CG
over frames, against the FP64 oracle's joint solvers.

3rd_dim_cart (NFFT.hpp:84-103, :331-412; NCSENSE.hpp:81-99, :338-339): Nz
2D slices on one trajectory and one set of weights, maps (N0, N1, Nz, C), data
(nk, Nz, C); codeare's CGLS runs once on the whole (N0, N1, Nz) image, and so
does a nmany frame series (NCSENSE.hpp:141-151, :251-348).  Tolerances
(relative L2): operators 1e-5, fixed-iteration solves 1e-4 (as the single-frame
parity tests)."""
import numpy as np
import pytest

from helpers import rel, rand_img

pytestmark = pytest.mark.gpu

OP_TOL = 1e-5


def _slices(cgs, ora, N=48, C=4, Nz=3, spokes=24, R=96, seed=21, sigma=2.0, m=3):
    from codeare_amd import data
    rng = np.random.default_rng(seed)
    k = data.radial_2d(spokes, R, True, 0)
    M = k.shape[1]
    w = data.density(k, "k", R)
    S = np.asfortranarray(np.stack([data.coil_maps(N, 2, C, rng) * (1.0 + 0.3 * z) for z in range(Nz)], axis=2)
                          .astype(np.complex64))
    x = rand_img(rng, (N, N, Nz))
    O = ora.NFFT((N, N), M, sigma, m)
    O.kspace(k)
    # codeare's slice data (nk, Nz, C)
    y = np.zeros((M, Nz, C), dtype=np.complex64, order="F")
    for z in range(Nz):
        y[:, z, :] = O.E(np.asfortranarray(S[:, :, z, :]), x[..., z])
    y += (0.05 * np.abs(y).max() * (rng.normal(size=y.shape) + 1j * rng.normal(size=y.shape))).astype(np.complex64)
    return k, w, S, x, y, O, M


def test_stack_of_stars_operators(cgs, ora):
    """E and the reference E^H of NCSENSE with 3rd_dim_cart, slice by slice
    with each slice's maps and intensity map."""
    k, w, S, x, y, O, M = _slices(cgs, ora)
    Nz, C = S.shape[2], S.shape[3]
    ncs = cgs.NCSENSE({"sensitivities": S, "nk": M, "m": 3, "alpha": 2.0, "3rd_dim_cart": True, "ftiter": 3,
                       "fteps": 1e-3})
    ncs.KSpace(k)
    ncs.Weights(w)
    yg = ncs * x
    assert yg.shape == (M, Nz, C)
    xg = ncs / y
    assert xg.shape == S.shape[:3]
    for z in range(Nz):
        Sz = np.asfortranarray(S[:, :, z, :])
        e1 = rel(yg[:, z, :], O.E(Sz, x[..., z]))
        e2 = rel(xg[..., z], O.EH_ref(Sz, w, np.asfortranarray(y[:, z, :]), ftiter=3, fteps=1e-3))
        print(f"slice {z}: E {e1:.2e}  E^H(ref) {e2:.2e}")
        assert e1 <= OP_TOL and e2 <= OP_TOL


@pytest.mark.parametrize("sigma,m", [(2.0, 3), (1.0, 1)])
def test_stack_of_stars_cgls(cgs, ora, sigma, m):
    """NCSENSE ->* with 3rd_dim_cart: one CGLS over the slices with the
    reference E^H (ftiter 3, squared-ratio cgeps), against the oracle's joint
    CGLS; sigma 1 / m 1 is what codeare's CGSENSE module passes (no alpha)."""
    k, w, S, x, y, O, M = _slices(cgs, ora, sigma=sigma, m=m, seed=22)
    Nz = S.shape[2]
    ncs = cgs.NCSENSE({"sensitivities": S, "nk": M, "m": m, "alpha": sigma, "3rd_dim_cart": True,
                       "cgiter": 5, "cgeps": 1e-7, "lambda": 5e-2, "ftiter": 3, "fteps": 1e-3})
    ncs.KSpace(k)
    ncs.Weights(w)
    xg = ncs.Adjoint(y)
    plans = []
    for z in range(Nz):
        Oz = ora.NFFT(O.N, M, sigma, m)
        Oz.kspace(k)
        plans.append(Oz)
    yb = np.concatenate([y[:, z, :].reshape(-1, order="F") for z in range(Nz)])
    xo, it, hist = ora.cgls_ref_frames(plans, S, np.tile(w, Nz), yb, mode=1, ftiter=3, fteps=1e-3, maxit=5,
                                       cgeps=1e-7, lam=5e-2, maps_per_frame=True)
    e = rel(xg, xo)
    rep = ncs.last_report
    print(f"sigma {sigma} m {m}: joint CGLS image {e:.2e}; iters gpu {rep['iters']} oracle {it}")
    assert all(i == it for i in rep["iters"])
    assert e <= 1e-4


def _frames(cgs, N=48, C=4, R=96, counts=(7, 13, 4), seed=23, per_frame_maps=False):
    from codeare_amd import data
    rng = np.random.default_rng(seed)
    ks, first = [], 0
    for c in counts:
        ks.append(data.radial_2d(c, R, True, first))
        first += c
    k = np.asfortranarray(np.concatenate(ks, axis=1))
    off = np.concatenate([[0], np.cumsum(np.array(counts) * R)]).astype(np.int64)
    w = data.density(k, "k", R)
    F = len(counts)
    if per_frame_maps:
        S = np.asfortranarray(np.stack([data.coil_maps(N, 2, C, rng) for _ in range(F)], axis=2))
    else:
        S = data.coil_maps(N, 2, C, rng)
    x = rand_img(rng, (N, N, F))
    P = cgs.Plan((N, N), k.shape[1], 2.0, 3, ncoils=C, frame_offsets=off, maps_per_frame=per_frame_maps)
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    return P, k, w, S, x, off


@pytest.mark.parametrize("per_frame_maps", [False, True])
def test_joint_true_cg_frames(cgs, ora, per_frame_maps):
    """Joint CG (true adjoint, FP64 scalars) over uneven frames, through the
    NUFFT operator and the Toeplitz one (kernels launched one by one and as a
    graph), against the oracle's joint CG."""
    P, k, w, S, x, off = _frames(cgs, per_frame_maps=per_frame_maps)
    C = S.shape[-1]
    y = P.E(x)
    F = len(off) - 1
    plans = []
    for f in range(F):
        Of = ora.NFFT(P.N, int(off[f + 1] - off[f]), 2.0, 3)
        Of.kspace(k[:, off[f]:off[f + 1]])
        plans.append(Of)
    xo, it, hist = ora.cg_true_frames(plans, S, w, y, maxit=10, tol=0.0, lam=0.05, maps_per_frame=per_frame_maps)
    for toep, gi in ((False, 0), (True, 0), (True, 4)):
        xg, rep = P.solve(y, maxit=10, tol=0.0, lam=0.05, toeplitz=toep, graph_iters=gi, joint=True)
        e = rel(xg, xo)
        print(f"toeplitz={toep} graph={gi}: joint CG(10) image {e:.2e}, iters {rep['iters']}, "
              f"rel {rep['final_rel'][0]:.3e} vs {hist[-1]:.3e}")
        assert rep["iters"] == [10] * F
        assert e <= 1e-4
        assert abs(rep["final_rel"][0] - hist[-1]) <= 1e-3 * hist[-1]
    # to convergence (lambda ~1e-2 of the largest eigenvalue, so that the
    # stopped images are close to the solution): one stop for all frames
    xo, it, hist = ora.cg_true_frames(plans, S, w, y, maxit=300, tol=1e-5, lam=2.0, maps_per_frame=per_frame_maps)
    for toep in (False, True):
        xg, rep = P.solve(y, maxit=300, tol=1e-5, lam=2.0, toeplitz=toep, graph_iters=4 if toep else 0, joint=True)
        e = rel(xg, xo)
        print(f"joint to 1e-5, toeplitz={toep}: iters gpu {rep['iters']} oracle {it}, image {e:.2e}")
        assert len(set(rep["iters"])) == 1 and abs(rep["iters"][0] - it) <= 1
        assert rep["final_rel"][0] <= 1e-5
        # the Toeplitz operator is the exact normal matrix of the NUDFT, the
        # oracle's the NFFT's (1e-5 apart per application): over ~100 iterations
        # the solutions drift apart by a few 1e-4
        assert e <= (1e-3 if toep else 1e-4)


def test_lean_toeplitz_per_frame_maps(cgs):
    """Per-frame maps on the lean Toeplitz passes (256^2, 16 coils, three
    slices): Toeplitz against the NUFFT normal operator, and 12 joint CG
    iterations through each operator."""
    from codeare_amd import data
    rng = np.random.default_rng(31)
    N, C, Nz, R = 256, 16, 3, 512
    k = data.radial_2d(64, R, True, 0)
    M = k.shape[1]
    w = data.density(k, "k", R)
    S = np.asfortranarray(np.stack([data.coil_maps(N, 2, C, rng) for _ in range(Nz)], axis=2))
    x = rand_img(rng, (N, N, Nz))
    P = cgs.Plan((N, N), M * Nz, 2.0, 3, ncoils=C, frame_offsets=np.arange(Nz + 1, dtype=np.int64) * M,
                 maps_per_frame=True)
    P.kspace(np.asfortranarray(np.tile(k, (1, Nz))))
    P.weights(np.tile(w, Nz))
    P.sensitivities(S)
    e = rel(P.normal(x, lam=0.5, toeplitz=True), P.normal(x, lam=0.5))
    y = P.E(x)
    xt, rt = P.solve(y, maxit=12, tol=0.0, lam=1.0, toeplitz=True, graph_iters=4, joint=True)
    xn, rn = P.solve(y, maxit=12, tol=0.0, lam=1.0, joint=True)
    e2 = rel(xt, xn)
    # slice z alone (its own maps) is a different CG: check the maps went to
    # the right slices through the operator instead
    for z in range(Nz):
        P1 = cgs.Plan((N, N), M, 2.0, 3, ncoils=C)
        P1.kspace(k)
        P1.weights(w)
        P1.sensitivities(np.asfortranarray(S[:, :, z, :]))
        ez = rel(P.normal(x, lam=0.5, toeplitz=True)[..., z], P1.normal(np.asfortranarray(x[..., z]), lam=0.5,
                                                                       toeplitz=True))
        print(f"slice {z}: per-frame-maps normal vs single plan {ez:.2e}")
        assert ez <= 1e-5
    print(f"lean per-frame maps: toeplitz vs nufft normal {e:.2e}, joint CG(12) {e2:.2e}")
    assert e <= 2e-5 and e2 <= 1e-4
    assert rt["iters"] == rn["iters"] == [12] * Nz


def test_nmany_joint_cgls(cgs, ora):
    """NCSENSE with dim4 frames (shared maps, per-frame trajectories, shared
    weights): codeare's one CGLS over all frames, against the oracle."""
    from codeare_amd import data
    rng = np.random.default_rng(24)
    N, C, R, F, spokes = 48, 4, 96, 3, 16
    ks = [data.radial_2d(spokes, R, True, f * spokes) for f in range(F)]
    M = ks[0].shape[1]
    k3 = np.asfortranarray(np.stack(ks, axis=2))          # (2, nk, dim4) as NCSENSE::KSpace takes it
    w = data.density(ks[0], "k", R)                         # one weight vector for every frame
    S = data.coil_maps(N, 2, C, rng)
    x = rand_img(rng, (N, N, F))
    ncs = cgs.NCSENSE({"sensitivities": S, "nk": M, "m": 3, "alpha": 2.0, "dim4": F, "cgiter": 5,
                       "cgeps": 1e-7, "lambda": 5e-2, "ftiter": 3, "fteps": 1e-3})
    ncs.KSpace(k3)
    ncs.Weights(w)
    y = ncs * x
    xg = ncs.Adjoint(y)
    plans = []
    for f in range(F):
        Of = ora.NFFT((N, N), M, 2.0, 3)
        Of.kspace(ks[f])
        plans.append(Of)
    xo, it, _ = ora.cgls_ref_frames(plans, S, np.tile(w, F), y, mode=1, ftiter=3, fteps=1e-3, maxit=5,
                                    cgeps=1e-7, lam=5e-2)
    e = rel(xg, xo)
    print(f"nmany joint CGLS image {e:.2e}; iters gpu {ncs.last_report['iters']} oracle {it}")
    assert all(i == it for i in ncs.last_report["iters"])
    assert e <= 1e-4


def test_nfft_3rd_dim_cart(cgs, ora):
    """NFFT<cxfl> with 3rd_dim_cart: Trafo and the iterative Adjoint slice by
    slice on one trajectory."""
    rng = np.random.default_rng(25)
    N, Nz, M = 32, 4, 2500
    k = np.asfortranarray(rng.uniform(-0.5, 0.5, size=(2, M)).astype(np.float32))
    w = rng.uniform(0.1, 1.0, M).astype(np.float32)
    ft = cgs.NFFT({"nk": M, "imsz": (N, N, Nz), "m": 3, "alpha": 2.0, "3rd_dim_cart": True, "maxit": 3,
                   "epsilon": 1e-12})
    ft.KSpace(k)
    ft.Weights(w)
    img = rand_img(rng, (N, N, Nz))
    f = ft * img
    assert f.shape == (M, Nz)
    back = ft.Adjoint(f)
    assert back.shape == (N, N, Nz)
    O = ora.NFFT((N, N), M, 2.0, 3)
    O.kspace(k)
    for z in range(Nz):
        e1 = rel(f[:, z], O.trafo(img[..., z]))
        o, it, _ = O.itrafo(f[:, z], w, maxit=3, eps=1e-12)
        e2 = rel(back[..., z], o)
        print(f"slice {z}: trafo {e1:.2e} adjoint {e2:.2e}")
        assert e1 <= OP_TOL and e2 <= OP_TOL


def test_empty_and_ragged_frames(cgs):
    """Frames of 7, 0, 13 and 1 spokes in one plan: the empty frame gets no
    data, a zero image from E^H and a zero solution after 0 iterations; the
    others match single-plan operators and solves (per-frame CG), and a joint
    solve equals the joint solve of the plan without the empty frame."""
    from codeare_amd import data
    rng = np.random.default_rng(41)
    N, C, R = 48, 4, 96
    counts = (7, 0, 13, 1)
    ks, first = [], 0
    for c in counts:
        ks.append(data.radial_2d(c, R, True, first))
        first += c
    k = np.asfortranarray(np.concatenate(ks, axis=1))
    off = np.concatenate([[0], np.cumsum(np.array(counts) * R)]).astype(np.int64)
    w = data.density(k, "k", R)
    S = data.coil_maps(N, 2, C, rng)
    x = rand_img(rng, (N, N, len(counts)))
    P = cgs.Plan((N, N), k.shape[1], 2.0, 3, ncoils=C, frame_offsets=off)
    P.kspace(k)
    P.weights(w)
    P.sensitivities(S)
    y = P.E(x)
    assert y.size == C * off[-1]
    xe = P.EH(y)
    assert np.all(xe[..., 1] == 0)
    for toep in (False, True):
        xs, rep = P.solve(y, maxit=200, tol=1e-4, lam=0.5, toeplitz=toep, graph_iters=4 if toep else 0)
        assert rep["iters"][1] == 0 and np.all(xs[..., 1] == 0)
        for f in (0, 2, 3):
            Mf = int(off[f + 1] - off[f])
            P1 = cgs.Plan((N, N), Mf, 2.0, 3, ncoils=C)
            P1.kspace(k[:, off[f]:off[f + 1]])
            P1.weights(w[off[f]:off[f + 1]])
            P1.sensitivities(S)
            yf = y[C * off[f]: C * off[f + 1]].reshape((Mf, C), order="F")
            assert rel(P1.E(x[..., f]), yf) <= 1e-6
            x1, r1 = P1.solve(yf, maxit=200, tol=1e-4, lam=0.5, toeplitz=toep)
            ef = rel(xs[..., f], x1)
            print(f"toeplitz={toep} frame {f} ({counts[f]} spokes): iters {rep['iters'][f]} vs {r1['iters'][0]}, "
                  f"image {ef:.2e}")
            assert abs(rep["iters"][f] - r1["iters"][0]) <= 1
            assert ef <= (1e-5 if rep["iters"][f] == r1["iters"][0] else 1e-3)
    # joint: the empty frame adds nothing to the sums
    xj, rj = P.solve(y, maxit=30, tol=0.0, lam=0.5, joint=True)
    keep = [0, 2, 3]
    off2 = np.concatenate([[0], np.cumsum(np.array([counts[f] for f in keep]) * R)]).astype(np.int64)
    k2 = np.asfortranarray(np.concatenate([ks[f] for f in keep], axis=1))
    P2 = cgs.Plan((N, N), k2.shape[1], 2.0, 3, ncoils=C, frame_offsets=off2)
    P2.kspace(k2)
    P2.weights(data.density(k2, "k", R))
    P2.sensitivities(S)
    y2 = np.concatenate([y[C * off[f]: C * off[f + 1]] for f in keep])
    x2, r2 = P2.solve(y2, maxit=30, tol=0.0, lam=0.5, joint=True)
    ej = rel(xj[..., keep], x2)
    print(f"joint with / without the empty frame: {ej:.2e}")
    assert ej <= 1e-5 and np.all(xj[..., 1] == 0)
