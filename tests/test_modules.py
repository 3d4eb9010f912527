"""codeare's .cod files and the CGSENSE / NuFFT module harnesses
(src/modules/CGSENSE.cpp, src/modules/NuFFT.cpp) over the GPU operators."""
import os
import struct

import numpy as np
import pytest

from codeare_amd import codfile
from helpers import rel


def test_cod_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    for a in [rng.normal(size=(3, 4)).astype(np.float32),
              (rng.normal(size=(5, 2, 3)) + 1j * rng.normal(size=(5, 2, 3))).astype(np.complex64),
              rng.normal(size=7), np.arange(6, dtype=np.int16).reshape(2, 3)]:
        p = str(tmp_path / "m.cod")
        codfile.write_cod(p, a, "signals", res=[0.5] * a.ndim)
        b, name, res = codfile.read_cod(p)
        assert name == "signals" and b.dtype == a.dtype and b.shape == a.shape
        assert np.array_equal(a, b) and res == (0.5,) * a.ndim


def test_cod_byte_layout(tmp_path):
    """A file assembled field by field as CODFile::Write emits it (CODFile.hpp:150-186)."""
    a = np.asfortranarray(np.array([[1 + 2j, 3 + 4j, 5 + 6j], [7 + 8j, 9 + 10j, 11 + 12j]], np.complex64))
    raw = struct.pack("<i", 2) + struct.pack("<Q", 2) + struct.pack("<2Q", 2, 3) + struct.pack("<2f", 1, 1)
    raw += struct.pack("<Q", 6) + b"kspace"
    raw += np.array([1 + 2j, 7 + 8j, 3 + 4j, 9 + 10j, 5 + 6j, 11 + 12j], np.complex64).tobytes()
    raw += b"543f562189f1e82beb9c177f89f67822"
    p = str(tmp_path / "k.cod")
    with open(p, "wb") as f:
        f.write(raw)
    b, name, _ = codfile.read_cod(p)
    assert name == "kspace" and np.array_equal(a, b)
    q = str(tmp_path / "w.cod")
    codfile.write_cod(q, a, "kspace")
    assert open(q, "rb").read() == raw


def test_cod_rejects_truncated(tmp_path):
    p = str(tmp_path / "t.cod")
    codfile.write_cod(p, np.ones((4, 4), np.float32), "w")
    data = open(p, "rb").read()
    with open(p, "wb") as f:
        f.write(data[:-40])
    with pytest.raises(ValueError):
        codfile.read_cod(p)


@pytest.mark.gpu
def test_cgsense_module_matches_oracle(ora, tmp_path):
    """The CGSENSE module from .cod files: NCSENSE with sigma 1 (no alpha
    passed), fteps 1e-3, the reference E^H and CGLS -- against the oracle's
    CGLS with the same settings."""
    from codeare_amd import data, modules
    prob = data.make_problem(1, 0, N=64, coils=4, spokes=48, samples=128)
    N, C, M = 64, 4, prob.k.shape[1]
    O = ora.NFFT((N, N), M, 1.0, 3)
    O.kspace(prob.k)
    y = data.make_data(prob, lambda x: O.E(prob.maps, x))
    files = {}
    for name, arr in [("sensitivities", prob.maps), ("kspace", prob.k), ("weights", prob.w), ("signals", y)]:
        files[name] = str(tmp_path / f"{name}.cod")
        codfile.write_cod(files[name], arr, name)
    out = str(tmp_path / "image.cod")
    rc = modules.main(["CGSENSE", f"nk={M}", "m=3", "lambda=5e-2", "cgeps=1e-7", "cgmaxit=6", "ftmaxit=3",
                       *[f"{k}={v}" for k, v in files.items()], f"image={out}"])
    assert rc == 0
    img, name, _ = codfile.read_cod(out)
    xo, it, _ = O.cgls_ref(prob.maps, prob.w, y, mode=1, ftiter=3, fteps=1e-3, maxit=6, cgeps=1e-7, lam=5e-2)
    e = rel(img, xo)
    print(f"CGSENSE module (sigma 1): image vs oracle {e:.2e}, oracle iterations {it}")
    assert name == "image" and img.shape == (N, N)
    assert e <= 1e-4


@pytest.mark.gpu
def test_nufft_module_matches_oracle(ora):
    from codeare_amd import modules
    from helpers import rand_traj, rand_img
    rng = np.random.default_rng(8)
    N, M = 32, 2000
    k = rand_traj(rng, 2, M)
    w = rng.uniform(0.2, 1.0, M).astype(np.float32)
    f = rand_img(rng, (M,))
    s = modules.NuFFT()
    s.init(dim=2, Nx=N, Ny=N, M=M, shots=1, maxit=50, epsilon=1e-9, m=3, alpha=2.0)
    ws = {"kspace": k, "weights": w, "data": f}
    s.prepare(ws)
    img = s.process(ws)
    O = ora.NFFT((N, N), M, 2.0, 3)
    O.kspace(k)
    io, it, _ = O.itrafo(f, w, maxit=3, eps=7e-4)   # NFFT defaults, not the attributes
    e = rel(img, io)
    print(f"NuFFT module: {e:.2e} (oracle ITrafo steps {it})")
    assert e <= 1e-5


@pytest.mark.gpu
def test_cgsense_module_stack_of_stars(ora, tmp_path):
    """The CGSENSE module with dft_3rd_dim from .cod files: (N0, N1, Nz, C)
    maps, (nk, Nz, C) signals, one CGLS over the slices -- against the
    oracle's joint CGLS (sigma 1, fteps 1e-3, as the module passes them)."""
    from codeare_amd import data, modules
    rng = np.random.default_rng(12)
    N, C, Nz, R = 32, 4, 3, 64
    k = data.radial_2d(32, R, True, 0)
    M = k.shape[1]
    w = data.density(k, "k", R)
    S = np.asfortranarray(np.stack([data.coil_maps(N, 2, C, rng) * (1 + 0.2 * z) for z in range(Nz)], axis=2)
                          .astype(np.complex64))
    x = (rng.normal(size=(N, N, Nz)) + 1j * rng.normal(size=(N, N, Nz))).astype(np.complex64)
    O = ora.NFFT((N, N), M, 1.0, 3)
    O.kspace(k)
    y = np.zeros((M, Nz, C), dtype=np.complex64, order="F")
    for z in range(Nz):
        y[:, z, :] = O.E(np.asfortranarray(S[:, :, z, :]), np.asfortranarray(x[..., z]))
    files = {}
    for name, arr in [("sensitivities", S), ("kspace", k), ("weights", w), ("signals", y)]:
        files[name] = str(tmp_path / f"{name}.cod")
        codfile.write_cod(files[name], arr, name)
    out = str(tmp_path / "image.cod")
    rc = modules.main(["CGSENSE", f"nk={M}", "m=3", "lambda=5e-2", "cgeps=1e-7", "cgmaxit=6", "ftmaxit=3",
                       "dft_3rd_dim=1", *[f"{k_}={v}" for k_, v in files.items()], f"image={out}"])
    assert rc == 0
    img, name, _ = codfile.read_cod(out)
    plans = []
    for z in range(Nz):
        Oz = ora.NFFT((N, N), M, 1.0, 3)
        Oz.kspace(k)
        plans.append(Oz)
    yb = np.concatenate([y[:, z, :].reshape(-1, order="F") for z in range(Nz)])
    xo, it, _ = ora.cgls_ref_frames(plans, S, np.tile(w, Nz), yb, mode=1, ftiter=3, fteps=1e-3, maxit=6,
                                    cgeps=1e-7, lam=5e-2, maps_per_frame=True)
    e = rel(img, xo)
    print(f"CGSENSE module, 3 slices: image vs oracle {e:.2e}, oracle iterations {it}")
    assert img.shape == (N, N, Nz)
    assert e <= 1e-4


@pytest.mark.gpu
def test_nufft_module_stack_of_stars(ora):
    from codeare_amd import modules
    from helpers import rand_traj, rand_img
    rng = np.random.default_rng(13)
    N, Nz, M = 32, 3, 1500
    k = rand_traj(rng, 2, M)
    w = rng.uniform(0.2, 1.0, M).astype(np.float32)
    f = rand_img(rng, (M, Nz))
    s = modules.NuFFT()
    s.init(dim=3, Nx=N, Ny=N, Nz=Nz, M=M, shots=1, maxit=50, epsilon=1e-9, m=3, alpha=2.0, dft_3rd_dim=1)
    ws = {"kspace": k, "weights": w, "data": f}
    s.prepare(ws)
    img = s.process(ws)
    assert img.shape == (N, N, Nz)
    O = ora.NFFT((N, N), M, 2.0, 3)
    O.kspace(k)
    for z in range(Nz):
        io, it, _ = O.itrafo(f[:, z], w, maxit=3, eps=7e-4)
        e = rel(img[..., z], io)
        print(f"NuFFT module slice {z}: {e:.2e}")
        assert e <= 1e-5


def test_estimate_sensitivities_spline_is_gsl_cspline():
    """interp1's default (GSL cspline, PolyVal.hpp:261-274) is the natural
    cubic spline: pinned against scipy's CubicSpline(bc_type="natural"), and
    exact on straight lines."""
    import scipy.interpolate as si
    from codeare_amd.modules import natural_cspline
    rng = np.random.default_rng(3)
    for n in (3, 5, 37):
        x = np.linspace(0.0, 1.0, n)
        y = rng.normal(size=(n, 2))
        xi = np.linspace(0.0, 1.0, 10 * n)
        assert np.abs(natural_cspline(x, y, xi) - si.CubicSpline(x, y, bc_type="natural")(xi)).max() <= 1e-12
    x = np.sort(rng.uniform(0, 2, 11))
    xi = np.linspace(x[0], x[-1], 57)
    assert np.abs(natural_cspline(x, 3 * x - 1, xi) - (3 * xi - 1)).max() <= 1e-12


def test_estimate_sensitivities_trajectory():
    """The module's k-space: "sync" (kx, ky interleaved in its first column)
    parsed to the raster spoke, interpolated 10x and cut to nk samples; volume
    v is the spoke turned by v golden angles (111.246... degrees); max |k| is
    1/2 (EstimateSensitivities.cpp:74-131)."""
    from codeare_amd import modules
    nr, nk, nv = 40, 256, 5
    t = np.linspace(-1.0, 1.0, nr)
    raster = np.stack([0.7 * t, 0.2 * t], axis=1)                # a straight spoke
    sync = np.zeros((2 * nr, 3), np.float32)
    sync[:, 0] = raster.reshape(-1)                              # kx0, ky0, kx1, ky1, ...
    s = modules.EstimateSensitivities.__new__(modules.EstimateSensitivities)
    s.prepare({"sync": sync})
    assert np.allclose(s.raster, raster, atol=1e-7)
    k = s.trajectory(nk, nv)
    assert k.shape == (2, nk * nv) and k.dtype == np.float32
    kc = (k[0] + 1j * k[1]).reshape((nk, nv), order="F")
    assert abs(np.abs(kc).max() - 0.5) <= 1e-6
    # the interpolated spoke: a line from -1 towards the first 10 % of the raster span
    XI = np.linspace(0.0, 1.0, 10 * nr)[:nk]
    line = (-1.0 + 2.0 * XI)
    ref0 = (0.7 * line + 1j * 0.2 * line)
    scale = 0.5 / np.abs(ref0).max()
    assert np.abs(kc[:, 0] - ref0 * scale).max() <= 1e-5
    for v in range(1, nv):
        turn = np.angle(np.vdot(kc[:, 0], kc[:, v]))
        want = np.angle(np.exp(1j * np.deg2rad(v * modules.NYUGA)))
        assert abs(np.angle(np.exp(1j * (turn - want)))) <= 1e-5


def test_estimate_sensitivities_rejects_mismatched_sizes():
    from codeare_amd import modules
    s = modules.EstimateSensitivities.__new__(modules.EstimateSensitivities)
    s.attrs = {"cart_3rd_dim": 0, "all_coils": 0}
    s.imsz = [16, 16]
    s.nk_attr = 100
    s.raster = np.zeros((10, 2), np.float32)
    with pytest.raises(ValueError):
        s.process({"meas": np.zeros((30, 2, 1, 3), np.complex64)})


@pytest.mark.gpu
def test_estimate_sensitivities_module_matches_oracle(ora):
    """Coil 0 of the module's output is the NFFT's iterative adjoint (ftiter
    CGNR steps, unit weights) of 1e6 x coil 0's data on the golden-angle
    trajectory -- against the oracle's ITrafo; the other coils stay zero
    (the module's `i < 1` loop) unless all_coils=1."""
    from codeare_amd import modules
    rng = np.random.default_rng(21)
    # two partitions (cart_3rd_dim): the module squeezes meas before it reads
    # (nk, nc, nl, nv), so a single-line acquisition would lose its nl axis
    N, nr, nk, nv, nc, nl = 32, 12, 64, 24, 3, 2
    t = np.linspace(-1.0, 1.0, nr)
    sync = np.zeros((2 * nr, 1), np.float32)
    sync[:, 0] = np.stack([t, 0.3 * t ** 3], axis=1).reshape(-1)
    meas = ((rng.normal(size=(nk, nc, nl, nv)) + 1j * rng.normal(size=(nk, nc, nl, nv))) * 1e-6).astype(np.complex64)
    for all_coils in (0, 1):
        s = modules.EstimateSensitivities()
        s.init(nk=nk, nl=nv, image_size=f"{N},{N},{nl}", ftalpha=2.0, ftm=3, ftiter=3, cart_3rd_dim=1,
               all_coils=all_coils)
        ws = {"meas": meas, "sync": sync}
        s.prepare(ws)
        sens = s.process(ws)
        assert sens.shape == (N, N, nl, nc)
        k = ws["kspace"]
        O = ora.NFFT((N, N), nk * nv, 2.0, 3)
        O.kspace(k)
        for c in range(nc):
            if c > 0 and not all_coils:
                assert not np.any(sens[..., c])
                continue
            for z in range(nl):
                # (nk, nv) samples of coil c, partition z, times 1e6
                f = np.asfortranarray(meas[:, c, z, :]).reshape(-1, order="F") * np.float32(1e6)
                io, it, _ = O.itrafo(f.astype(np.complex64), np.ones(nk * nv, np.float32), maxit=3, eps=7e-4)
                e = rel(sens[:, :, z, c], io)
                print(f"EstimateSensitivities all_coils={all_coils} coil {c} partition {z}: {e:.2e} "
                      f"(oracle steps {it})")
                assert e <= 1e-5


def test_codeare_phantom():
    """codeare's phantom<cxfl>(n) (Creators.hpp:369-489): the centre lies in the
    two outer ellipses only (1 - 0.8); the bright ring between them covers the
    area the two ellipses' semi-axes give; the 3D phantom keeps codeare's
    ellipsoid quirk (the slice index where the value belongs)."""
    from codeare_amd.modules import phantom
    n = 512
    P = phantom(n)
    assert P.shape == (n, n) and P.dtype == np.complex64 and not np.any(P.imag)
    assert abs(P[n // 2, n // 2] - 0.2) < 1e-6
    ring = np.count_nonzero(np.isclose(P.real, 1.0))
    want = np.pi * (0.69 * 0.92 - 0.6624 * 0.874) * n * n / 4   # the second ellipse sits 0.0184 off centre
    assert abs(ring - want) / want < 0.08
    # the dark ellipses (-0.2, centres offset -/+0.22 along the second index:
    # m0 = (1 + 0.22) n / 2 = 312, (1 - 0.22) n / 2 = 200) hold 0; the same
    # offsets along the first index fall in the 0.2 interior (pins which index is which)
    assert abs(P[256, 312]) < 1e-6 and abs(P[256, 200]) < 1e-6
    assert abs(P[312, 256] - 0.2) < 1e-6 and P[200, 256].real > 0.1
    Q = phantom(32, 3)
    assert Q.shape == (32, 32, 32)
    # voxel (16, 16, 16): inside ellipsoids 1 and 2 (offset 0), so 2 x its slice index
    assert abs(Q[16, 16, 16] - 32.0) < 1e-6
    # every value inside is a multiple of its slice index
    s = np.arange(32)[None, None, :]
    k = np.divide(Q.real, s, out=np.zeros_like(Q.real), where=s > 0)
    assert np.allclose(k, np.round(k))


@pytest.mark.gpu
def test_module_testcases_match_oracle(ora):
    """CGSENSE testcase=1 (CGSENSE.cpp:171-172) and NuFFT test_case=1
    (NuFFT.cpp:138-140) take E / the NFFT of codeare's Shepp-Logan phantom as
    their data; against the oracle on the same phantom."""
    from codeare_amd import data, modules
    from codeare_amd.modules import phantom
    rng = np.random.default_rng(5)
    N, C, R = 64, 4, 128
    k = data.radial_2d(48, R, True)
    M = k.shape[1]
    w = data.density(k, "k", R)
    S = data.coil_maps(N, 2, C, rng)
    s = modules.CGSENSE()
    s.init(nk=M, m=3, testcase=1, cgmaxit=6, cgeps=1e-7, ftmaxit=3, **{"lambda": 5e-2})
    ws = {"sensitivities": S, "kspace": k, "weights": w}
    s.prepare(ws)
    img = s.process(ws)
    O = ora.NFFT((N, N), M, 1.0, 3)
    O.kspace(k)
    y = O.E(S, phantom(N))
    xo, it, _ = O.cgls_ref(S, w, y, mode=1, ftiter=3, fteps=1e-3, maxit=6, cgeps=1e-7, lam=5e-2)
    e = rel(img, xo)
    print(f"CGSENSE testcase: {e:.2e} vs oracle ({it} iterations)")
    assert e <= 1e-4
    u = modules.NuFFT()
    u.init(dim=2, Nx=N, Ny=N, M=M, shots=1, m=3, alpha=2.0, test_case=1)
    ws = {"kspace": k, "weights": w}
    u.prepare(ws)
    img = u.process(ws)
    O2 = ora.NFFT((N, N), M, 2.0, 3)
    O2.kspace(k)
    io, it, _ = O2.itrafo(O2.trafo(phantom(N)).astype(np.complex64), w, maxit=3, eps=7e-4)
    e = rel(img, io)
    print(f"NuFFT test_case: {e:.2e} vs oracle ITrafo")
    assert e <= 1e-5
