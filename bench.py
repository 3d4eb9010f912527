#!/usr/bin/env python3
"""CG-SENSE benchmark on one MI355X.

A step is one complete CG-SENSE solve of the 2D golden-angle workload
(256^2 image, 32 coils, 402 spokes x 512 samples, sigma 2, m 3, lambda 1e-3)
from zero to relative residual ||r||/||E^H W y|| <= 1e-5, with the
density-weighted true adjoint, through the C ABI (device-resident inputs).
Prints one JSON line with solves/s, the roofline of the dominant kernel
(HIP-event timed inside the timed region) and a CPU baseline (the FP64 C
port of the reference algorithm, timed on a bounded sample).
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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2, help="CG iterations in the CPU sample")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from codeare_amd import cgsense, data

    wl_id = args.workload
    prob = data.make_problem(wl_id, 0)
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

    maxit, tol, lam = 200, (wl.tol if wl.tol > 0 else 1e-5), wl.lam

    def step():
        _, rep = plan.solve(y, maxit=maxit, tol=tol, lam=lam, out=x)
        return rep

    for _ in range(args.warmup):
        rep = step()
    plan.kernel_stats(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        rep = step()
        iters.append(rep["max_iters"])
    plan.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = plan.kernel_stats(reset=True)
    ms_per_step = 1e3 * elapsed / args.steps
    solves_per_s = args.steps / elapsed

    # image error against the noiseless phantom, for the record
    xr = x.cpu().numpy()
    err_phantom = float(np.linalg.norm(xr - x_true.cpu().numpy()) / np.linalg.norm(x_true.cpu().numpy()))

    # roofline of the dominant kernel class (algorithmic bytes / HIP-event time)
    per = {}
    for k, s in stats.items():
        if s["launches"] > 0:
            per[k] = {"launches": s["launches"], "ms_total": s["ms"],
                      "us_per_launch": 1e3 * s["ms"] / s["launches"],
                      "GBps": (s["bytes"] / (s["ms"] * 1e-3)) / 1e9 if s["ms"] > 0 else None}
    dom = max(per, key=lambda k: per[k]["ms_total"])
    ach = per[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "us_per_launch": round(per[dom]["us_per_launch"], 2)}

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex vectors, fp64 CG scalars)",
        "data": "synthetic (seeded ellipse phantom, Gaussian coil maps, golden-angle radial)",
        "config": {"workload": wl.name, "image": list(N), "coils": C, "spokes": wl.spokes,
                   "samples_per_spoke": wl.samples, "M": M, "sigma": wl.sigma, "m": wl.m,
                   "lambda": lam, "tol": tol, "tol_kind": "norm ratio ||r||/||E^H W y||",
                   "eh": "density-weighted true adjoint", "start": "zero (plain)"},
        "cg_iters": iters,
        "rel_err_vs_phantom": round(err_phantom, 5),
        "plan_setup_s": round(t_plan, 3),
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in per.items()},
        "nufft_GBps": {"interp": per.get("interp", {}).get("GBps"), "grid": per.get("grid", {}).get("GBps")},
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, y_noisy, iters[-1], lam, args.cpu_iters)
        out["speedup_vs_cpu"] = round(out["value"] / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


def cpu_baseline(prob, y, gpu_iters, lam, cpu_iters):
    """The oracle's FP64 CG on the same inputs (the port of the reference
    algorithm: NFFT3-style FP64 NFFT, complex64 CG vectors), timed on a bounded
    sample: E^H y plus `cpu_iters` CG iterations, scaled to one solve of
    gpu_iters iterations."""
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
    return {"value": round(1.0 / t_solve, 5), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"E^H y + {cpu_iters} CG iterations of the same workload timed "
                      f"({t_b:.2f} s + {cpu_iters}x{t_iter:.2f} s), scaled to {gpu_iters} iterations; "
                      f"trajectory setup {t_setup:.2f} s excluded",
            "s_per_iter": round(t_iter, 3)}


if __name__ == "__main__":
    main()
