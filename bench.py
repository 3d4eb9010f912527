#!/usr/bin/env python3
"""CG-SENSE benchmark on one MI355X.

A step is one complete CG-SENSE solve of the 2D golden-angle workload
(256^2 image, 32 coils, 402 spokes x 512 samples, sigma 2, m 3, lambda 1e-3,
w = max(|k|, 1/512)) from zero to relative residual ||r||/||E^H W y|| <= 1e-5,
through the C ABI with inputs resident in HBM: E^H W y by gridding + FFT, then
CG iterations whose operator E^H W E + lambda is applied through the Toeplitz
embedding (the same operator as E then E^H, to NFFT accuracy).

The JSON line also carries: the same solve with the operator applied through
E and E^H (window gather, FFTs, gridding scatter), the HIP-event roofline of
the dominant kernel in the timed region, the NuFFT kernels' GB/s, the same
problem with lambda scaled to 1e-3 of the operator norm, and a CPU baseline
(the FP64 C port of the algorithm on the host cores, bounded sample, scaled).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CG-SENSE solves/s to rel-resid 1e-5, warm-start vs plain; NuFFT GB/s vs HBM"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--maxit", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--cpu-iters", type=int, default=2, help="CG iterations in the CPU sample")
    return ap.parse_args()


def per_kernel(stats):
    per = {}
    for k, s in stats.items():
        if s["launches"] > 0:
            per[k] = {"launches": s["launches"], "ms_total": round(s["ms"], 3),
                      "us_per_launch": round(1e3 * s["ms"] / s["launches"], 2),
                      "GBps": round((s["bytes"] / (s["ms"] * 1e-3)) / 1e9, 1) if s["ms"] > 0 else None}
    return per


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from codeare_amd import cgsense, data

    prob = data.make_problem(args.workload, 0)
    wl = prob.wl
    M = prob.k.shape[1]
    N = (wl.N,) * wl.d
    C = wl.coils
    plan = cgsense.Plan(N, M, wl.sigma, wl.m, ncoils=C, frame_offsets=prob.frame_offsets,
                        device_io=True, profile=True)
    t0 = time.perf_counter()
    plan.kspace(prob.k)
    plan.weights(prob.w)
    plan.sensitivities(prob.maps)
    t_plan = time.perf_counter() - t0

    dev = torch.device("cuda:0")
    x_true = torch.from_numpy(np.ascontiguousarray(prob.image.reshape(-1, order="F"))).to(dev)
    y = torch.empty(M * C, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    plan.E(x_true, out=y)
    plan.sync()
    y_h = y.cpu().numpy().reshape((M, C), order="F")
    y_noisy = data.add_noise(y_h, wl.snr_db, prob.rng_state)
    y = torch.from_numpy(np.ascontiguousarray(y_noisy.reshape(-1, order="F"))).to(dev)
    x = torch.empty(plan.Ntot * plan.F, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()

    tol = wl.tol if wl.tol > 0 else 1e-5
    lam = wl.lam

    def solve(toeplitz, lam_=lam):
        _, rep = plan.solve(y, maxit=args.maxit, tol=tol, lam=lam_, out=x, toeplitz=toeplitz)
        return rep

    # ---- timed region: plain (zero-start) CG-SENSE solves ----
    for _ in range(args.warmup):
        solve(True)
    plan.kernel_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = [solve(True) for _ in range(args.steps)]
    plan.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per = per_kernel(plan.kernel_stats(reset=True))
    ms_per_step = 1e3 * elapsed / args.steps
    solves_per_s = args.steps / elapsed
    iters = [r["max_iters"] for r in reps]
    final_rel = reps[-1]["final_rel"][0]
    xr = x.cpu().numpy()
    xt = x_true.cpu().numpy()
    err_phantom = float(np.linalg.norm(xr - xt) / np.linalg.norm(xt))

    dom = max(per, key=lambda k: per[k]["ms_total"])
    ach = per[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "us_per_launch": per[dom]["us_per_launch"]}

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex vectors, fp64 CG scalars)",
        "data": "synthetic (seeded ellipse phantom, Gaussian coil maps, golden-angle radial, 40 dB noise)",
        "config": {"workload": wl.name, "image": list(N), "coils": C, "spokes": wl.spokes,
                   "samples_per_spoke": wl.samples, "M": M, "sigma": wl.sigma, "m": wl.m,
                   "lambda": lam, "tol": tol, "tol_kind": "norm ratio ||r||/||E^H W y||",
                   "eh": "density-weighted true adjoint", "start": "zero (plain)",
                   "normal_op": "toeplitz embedding (E^H W y by gridding)"},
        "cg_iters": iters,
        "final_rel_resid": final_rel,
        "rel_err_vs_phantom": round(err_phantom, 5),
        "plan_setup_s": round(t_plan, 3),
        "roofline": roofline,
        "kernels": per,
    }

    if not args.no_extra:
        # same solve with E then E^H in every iteration (gather, FFTs, scatter)
        solve(False)
        plan.kernel_stats(reset=True)
        t0 = time.perf_counter()
        rep = solve(False)
        t_nufft = time.perf_counter() - t0
        pn = per_kernel(plan.kernel_stats(reset=True))
        out["nufft_operator_solve"] = {"solves_per_s": round(1.0 / t_nufft, 4), "iters": rep["max_iters"],
                                       "final_rel_resid": rep["final_rel"][0],
                                       "ms_per_iter": round(1e3 * t_nufft / max(1, rep["max_iters"]), 4),
                                       "kernels": pn}
        out["nufft_GBps"] = {k: pn[k]["GBps"] for k in ("interp", "grid", "fft_fwd", "fft_adj") if k in pn}
        out["nufft_frac_hbm"] = {k: round(v / HBM_PEAK_GBS, 4) for k, v in out["nufft_GBps"].items() if v}
        # the same problem with lambda at 1e-3 of the operator norm (power iteration)
        xs = torch.randn(plan.Ntot * plan.F, dtype=torch.complex64, device=dev)
        qv = torch.empty_like(xs)
        lmax = 0.0
        for _ in range(20):
            plan.normal(xs, out=qv, toeplitz=True)
            plan.sync()
            lmax = float(torch.linalg.vector_norm(qv) / torch.linalg.vector_norm(xs))
            xs = qv / torch.linalg.vector_norm(qv)
        lam_s = 1e-3 * lmax
        solve(True, lam_s)
        t0 = time.perf_counter()
        rep = solve(True, lam_s)
        t_s = time.perf_counter() - t0
        xr = x.cpu().numpy()
        out["scaled_lambda_solve"] = {"lambda": round(lam_s, 3), "lambda_max": round(lmax, 1),
                                      "solves_per_s": round(1.0 / t_s, 3), "iters": rep["max_iters"],
                                      "final_rel_resid": rep["final_rel"][0],
                                      "rel_err_vs_phantom": round(float(np.linalg.norm(xr - xt) / np.linalg.norm(xt)), 5)}

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, y_noisy, iters[-1], lam, args.cpu_iters)
        out["speedup_vs_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


def cpu_baseline(prob, y, gpu_iters, lam, cpu_iters):
    """The oracle's FP64 CG on the same inputs (the port of the reference
    algorithm: NFFT3-style FP64 NFFT with precomputed window, complex64 CG
    vectors, one OpenMP thread per coil), timed on a bounded sample: E^H y plus
    `cpu_iters` CG iterations, scaled to one solve of gpu_iters iterations."""
    from oracle import ora
    ora.build()
    threads = max(1, min(os.cpu_count() or 1, 16))
    ora.set_threads(threads)
    wl = prob.wl
    O = ora.NFFT((wl.N,) * wl.d, prob.k.shape[1], wl.sigma, wl.m)
    t0 = time.perf_counter()
    O.kspace(prob.k)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.cg_true(prob.maps, prob.w, y, maxit=0, tol=0.0, lam=lam)
    t_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.cg_true(prob.maps, prob.w, y, maxit=cpu_iters, tol=0.0, lam=lam)
    t_k = time.perf_counter() - t0
    t_iter = max(1e-9, (t_k - t_b) / cpu_iters)
    t_solve = t_b + gpu_iters * t_iter
    return {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"E^H y + {cpu_iters} CG iterations of the same workload timed "
                      f"({t_b:.2f} s + {cpu_iters}x{t_iter:.2f} s), scaled to the {gpu_iters} iterations the "
                      f"GPU solve needed; trajectory setup {t_setup:.2f} s excluded",
            "s_per_iter": round(t_iter, 3)}


if __name__ == "__main__":
    main()
