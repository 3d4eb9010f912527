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
problem with lambda scaled to 1e-3 of the operator norm, codeare's own
algorithm (CGLS with the reference E^H) on the GPU, a CPU baseline (the FP64 C
port on the host cores, bounded sample, scaled) in both modes, and the error of
the GPU images against the FP64 direct-NUDFT CG at equal iteration counts.
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
    ap.add_argument("--graph-iters", type=int, default=32,
                    help="CG iterations per HIP graph launch in the headline solves (0: none); the short "
                         "secondary solves use at most 8")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--cpu-iters", type=int, default=10,
                    help="timed CG iterations in the CPU sample (plus one warm-up); the median is scaled")
    ap.add_argument("--ref-maxit", type=int, default=3000,
                    help="iteration cap of the reference-algorithm (ref-compat CGLS) solve")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the 64-frame batched series")
    ap.add_argument("--volume", dest="volume", action="store_true", default=None,
                    help="time the 3D koosh-ball workload (default: with the secondary measurements)")
    ap.add_argument("--no-volume", dest="volume", action="store_false", help="skip the 3D koosh-ball workload")
    ap.add_argument("--reference-only", action="store_true",
                    help="only codeare's own algorithm on the headline data (CGLS with the reference E^H), one JSON line")
    ap.add_argument("--nufft-only", action="store_true",
                    help="only time the NuFFT operators (E, then E^H W) at 256^2 x 32 coils and the 128^3 x 16-coil "
                         "koosh ball, `steps` applications each (for kernel profiles)")
    ap.add_argument("--variants", action="store_true",
                    help="with --nufft-only: also run every alternative gridding / gather kernel (one plan each)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with 3 device copies of 256 MiB (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--warm-seeds", type=int, default=16, help="held-out seeds the warm-start leg evaluates")
    ap.add_argument("--weights", default=os.path.join(ROOT, "codeare_amd", "weights", "unet_golden2d_256_c32.npz"),
                    help="trained warm-start network (skipped if absent)")
    return ap.parse_args()


def per_kernel(stats):
    per = {}
    for k, s in stats.items():
        if s["launches"] > 0 and s["ms"] > 0:
            per[k] = {"launches": s["launches"], "ms_total": round(s["ms"], 3),
                      "us_per_launch": round(1e3 * s["ms"] / s["launches"], 2),
                      "bytes_per_launch": s["bytes"] / s["launches"],
                      "GBps": round((s["bytes"] / (s["ms"] * 1e-3)) / 1e9, 1)}
    return per


def traffic_for(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None if that workload/kernel has not been counted."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t[workload][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(sharding the path over several GPUs is not implemented)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from codeare_amd import cgsense, data

    if args.nufft_only:
        out = nufft_leg(args)
        if args.copy_calibration:
            copy_calibration(out)
        print(json.dumps(out))
        return
    prob = data.make_problem(args.workload, 0)
    wl = prob.wl
    M = prob.k.shape[1]
    N = (wl.N,) * wl.d
    C = wl.coils
    plan = cgsense.Plan(N, M, wl.sigma, wl.m, ncoils=C, frame_offsets=prob.frame_offsets,
                        device_io=True, profile=True)
    # torch's tensor work and the plans' kernels share one stream, so data
    # written by one side is ordered before the other side reads it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    plan.set_stream(stream.cuda_stream)
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
    if args.reference_only:
        print(json.dumps({"workload": wl.name, "reference_algorithm": {"gpu": ref_compat_leg(args, plan, y, x, lam)}}))
        return

    def solve(toeplitz, lam_=lam, graph=args.graph_iters):
        _, rep = plan.solve(y, maxit=args.maxit, tol=tol, lam=lam_, out=x, toeplitz=toeplitz,
                            graph_iters=graph)
        return rep

    # ---- timed region: plain (zero-start) CG-SENSE solves, K iterations per HIP graph ----
    plan.profile_mask([])          # launches and bytes counted, no timing events
    for _ in range(args.warmup):
        solve(True)
    plan.kernel_stats(reset=True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)                 # the plan's stream
    reps = [solve(True) for _ in range(args.steps)]
    ev1.record(stream)
    plan.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ev_ms_per_step = ev0.elapsed_time(ev1) / args.steps
    plan.kernel_stats(reset=True)
    ms_per_step = 1e3 * elapsed / args.steps
    solves_per_s = args.steps / elapsed
    iters = [r["max_iters"] for r in reps]
    final_rel = reps[-1]["final_rel"][0]
    xr = x.cpu().numpy()
    x_head = xr.copy()
    xt = x_true.cpu().numpy()
    err_phantom = float(np.linalg.norm(xr - xt) / np.linalg.norm(xt))

    # ---- kernel durations with HIP events on the plan's stream: one solve with
    # every class timed (the breakdown), then one with only the dominant class
    # timed (its average launch duration, least perturbed) ----
    # The dominant kernel's average launch duration in the timed solves: the
    # timed region's HIP-event time per CG iteration (the once-per-solve work
    # taken out) times the kernel's share of the iteration's kernel time.  The
    # share comes from one solve with every class timed by events around its
    # launches; those events add ~2 us to every launch, which cancels in a share
    # but not in a per-launch average (reported beside it, with the same
    # kernel's duration back to back on a scratch buffer).
    plan.profile_mask(None)
    solve(True, graph=0)
    per = per_kernel(plan.kernel_stats(reset=True))
    it_cls = [k for k, v in per.items() if v["launches"] >= iters[-1]]   # launched once per iteration
    if it_cls:
        dom = max(it_cls, key=lambda k: per[k]["ms_total"])
        share = per[dom]["ms_total"] / sum(per[k]["ms_total"] for k in it_cls)
        once_ms = sum(v["ms_total"] for k, v in per.items() if k not in it_cls)
        us_iter = (ev_ms_per_step - once_ms) * 1e3 / iters[-1]
        us = us_iter * share
    else:
        dom = max(per, key=lambda k: per[k]["ms_total"])
        share, us_iter, us = None, None, per[dom]["us_per_launch"]
    bpl = per[dom]["bytes_per_launch"]
    ach = round(bpl / (us * 1e-6) / 1e9, 1)
    aux = {"us_per_launch_events_per_launch": per[dom]["us_per_launch"]}
    if dom == "toep_conv":
        aux["us_per_launch_back_to_back"] = round(plan.time_toeplitz_conv(reps=200), 3)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_for(wl.name, dom),
                "us_per_launch": round(us, 3),
                "timing": "timed region's HIP events per CG iteration x the kernel's share of the iteration",
                "us_per_iteration": round(us_iter, 3) if us_iter else None,
                "share_of_iteration": round(share, 4) if share else None, **aux,
                "launches": per[dom]["launches"], "bytes_per_launch": round(bpl)}

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
        "graph_iters": args.graph_iters,
        "kernels": per,
    }

    if not args.no_extra:
        # same solve with E then E^H in every iteration (gather, FFTs, scatter)
        plan.profile_mask([])
        solve(False)
        t0 = time.perf_counter()
        rep = solve(False)
        t_nufft = time.perf_counter() - t0
        plan.profile_mask(None)
        plan.kernel_stats(reset=True)
        solve(False, graph=0)
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
        plan.profile_mask([])
        solve(True, lam_s, graph=min(args.graph_iters, 8))
        t0 = time.perf_counter()
        rep = solve(True, lam_s, graph=min(args.graph_iters, 8))
        t_s = time.perf_counter() - t0
        plan.profile_mask(None)
        xr = x.cpu().numpy()
        out["scaled_lambda_solve"] = {"lambda": round(lam_s, 3), "lambda_max": round(lmax, 1),
                                      "solves_per_s": round(1.0 / t_s, 3), "iters": rep["max_iters"],
                                      "final_rel_resid": rep["final_rel"][0],
                                      "rel_err_vs_phantom": round(float(np.linalg.norm(xr - xt) / np.linalg.norm(xt)), 5)}
        scaled = (lam_s, xr.copy(), rep["max_iters"])
        out["reference_algorithm"] = {"gpu": ref_compat_leg(args, plan, y, x, lam)}

        out["warm_start_solve"] = warm_start_leg(args, plan, y, x, tol, xt)
        if not args.no_dynamic:
            out["dynamic_batch"] = dynamic_leg(args)
    if args.volume if args.volume is not None else not args.no_extra:
        out["volume_3d"] = volume_leg(args)

    if not args.no_cpu:
        ref = out.get("reference_algorithm", {}).get("gpu")
        cb = cpu_baseline(prob, y_noisy, iters[-1], lam, args.cpu_iters,
                          ref_iters_gpu=ref["iters"] if ref else None)
        cref = cb.pop("reference_algorithm", None)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu_same_mode"] = round(out["value"] / cb["value"], 1)
        if ref and cref:
            out["reference_algorithm"]["cpu"] = cref
            out["reference_algorithm"]["gpu_vs_cpu"] = round(ref["solves_per_s"] / cref["solves_per_s"], 1)
        sc = scaled if not args.no_extra else (None, None, None)
        out["error_vs_fp64_nudft"] = fp64_errors(prob, y_noisy, x_head, iters[-1], lam, sc[0], sc[1], sc[2])
    if args.copy_calibration:
        copy_calibration(out)
    print(json.dumps(out))


def copy_calibration(out):
    """Three device copies of 256 MiB: the reference dispatch that scales the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_calibration_bytes"] = 2 * src.numel() * 4


# alternative kernels of the NuFFT operators, selected by the plan's
# environment switches at plan creation: (label, workload, switches)
NUFFT_VARIANTS = [
    ("2d_grid_bin", 2, {"CGS_GRID_BIN": "1"}),
    ("2d_grid_bin_cg4", 2, {"CGS_GRID_BIN": "1", "CGS_BIN_CG": "4", "CGS_BIN_THREADS": "512"}),
    ("3d_grid_bin", 3, {"CGS_GRID_BIN": "1"}),
    ("2d_grid_sweep", 2, {"CGS_GRID_SWEEP": "1"}),
    ("2d_grid_reg", 2, {"CGS_GRID_MFMA": "0"}),
    ("2d_grid_batched", 2, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0"}),
    ("2d_grid_atomic", 2, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0", "CGS_GRID_ATOMIC": "1"}),
    ("2d_grid_plain", 2, {"CGS_GRID_MFMA": "0", "CGS_GRID_REG": "0", "CGS_GRID_BATCHED": "0"}),
    ("2d_interp_global", 2, {"CGS_INTERP_LDS": "0"}),
    ("2d_interp_rows", 2, {"CGS_INTERP_CS": "0"}),
    ("2d_interp_x512", 2, {"CGS_INTERP_CS": "2"}),
    ("2d_interp_cs256", 2, {"CGS_INTERP_CS": "1"}),
    ("2d_interp_pairs256", 2, {"CGS_INTERP_CS": "3"}),
    ("2d_interp_cg32", 2, {"CGS_INTERP_CG": "32"}),
    ("2d_interp_cg8", 2, {"CGS_INTERP_CG": "8"}),
    ("2d_interp_item256", 2, {"CGS_TILE_ITEM_MAX": "256"}),
    ("2d_grid_vpi256", 2, {"CGS_VISITS_PER_ITEM": "256"}),
    ("3d_grid_reg3", 3, {"CGS_GRID_MFMA": "0"}),
    ("3d_interp_global", 3, {"CGS_INTERP_LDS": "0"}),
]


def nufft_leg(args):
    """E and E^H W (window gather / gridding scatter with their FFT passes) at
    256^2 x 32 coils (402 x 512 golden angle) and the 128^3 x 16-coil koosh
    ball, `steps` applications each after `warmup`; per-kernel HIP-event
    times and algorithmic GB/s."""
    import numpy as np
    import torch
    from codeare_amd import cgsense, data
    out = {"metric": "NuFFT operator GB/s (algorithmic) vs HBM", "unit": "GB/s"}
    runs = [(None, 2, {}), (None, 3, {})] + (NUFFT_VARIANTS if args.variants else [])
    probs = {}
    for label, cfg, env in runs:
        if cfg not in probs:
            probs[cfg] = data.make_problem(cfg, 0)
        prob = probs[cfg]
        wl = prob.wl
        M = prob.k.shape[1]
        saved = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            plan = cgsense.Plan((wl.N,) * wl.d, M, wl.sigma, wl.m, ncoils=wl.coils, device_io=True, profile=True)
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        print(f"nufft leg: {label or 'default'} ({wl.name})", file=sys.stderr, flush=True)
        plan.kspace(prob.k)
        plan.weights(prob.w)
        plan.sensitivities(prob.maps)
        dev = torch.device("cuda:0")
        x = torch.from_numpy(np.ascontiguousarray(prob.image.reshape(-1, order="F"))).to(dev)
        y = torch.empty(M * wl.coils, dtype=torch.complex64, device=dev)
        xb = torch.empty_like(x)
        for _ in range(max(1, args.warmup)):
            plan.E(x, out=y)
            plan.EH(y, out=xb)
        plan.sync()
        plan.kernel_stats(reset=True)
        for _ in range(args.steps):
            plan.E(x, out=y)
            plan.EH(y, out=xb)
        plan.sync()
        per = per_kernel(plan.kernel_stats(reset=True))
        out[label or wl.name] = {"applications": args.steps, "kernels": per,
                                 "frac_hbm": {k: round(v["GBps"] / HBM_PEAK_GBS, 4) for k, v in per.items()}}
        if label is None and args.variants:
            # what the autotuner picks on this trajectory, and its timings
            rep = plan.autotune(reps=5)
            out[wl.name]["autotune"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.items()}
        del plan
        torch.cuda.empty_cache()
    return out


def ref_compat_leg(args, plan, y, x, lam, ftiter=3, fteps=1e-3, cgeps=1e-10):
    """codeare's own solve on the GPU: CGLS (CGLS.hpp:30-59) with the reference
    E^H (ftiter steps of the NFFT3 weighted, damped CGNR per coil, then conj(S),
    coil sum and intensity map, NCSENSE.hpp:251-301), squared-ratio stop at
    cgeps = 1e-10 (= rel-resid 1e-5 as a norm ratio), same lambda and data as
    the headline; kernels launched one by one (no graph)."""
    from codeare_amd import cgsense

    def run(maxit):
        _, rep = plan.solve(y, maxit=maxit, tol=cgeps, tol_kind=cgsense.CGS_TOL_SQUARED_RATIO, lam=lam,
                            eh_mode=cgsense.CGS_EH_REF, ftiter=ftiter, fteps=fteps, out=x, check_nan=False)
        return rep

    plan.profile_mask([])
    run(3)
    plan.sync()
    t0 = time.perf_counter()
    rep = run(args.ref_maxit)
    plan.sync()
    el = time.perf_counter() - t0
    plan.kernel_stats(reset=True)
    return {"algorithm": f"CGLS, reference E^H (ftiter {ftiter}, fteps {fteps:g}), squared-ratio cgeps {cgeps:g}",
            "iters": rep["max_iters"], "final_squared_ratio": rep["final_rel"][0],
            "converged": bool(rep["final_rel"][0] <= cgeps), "maxit": args.ref_maxit,
            "solves_per_s": round(1.0 / el, 4), "ms_per_iter": round(1e3 * el / max(1, rep["max_iters"]), 4)}


def warm_start_leg(args, plan, y, x, tol, xt):
    """CG from the U-Net's output vs CG from zero on the same y with the same
    stopping test, at the lambda the network was trained for (1e-3 of
    lambda_max, stored with the weights: the literal lambda = 1e-3 is
    effectively unregularised).  The warm path's time includes E^H W y, the
    first steepest-descent scaling and the network inference."""
    import numpy as np
    import torch
    if not os.path.exists(args.weights):
        return {"skipped": f"no trained network at {os.path.relpath(args.weights, ROOT)}"}
    from codeare_amd.warmstart import UNet
    z = np.load(args.weights, allow_pickle=False)
    lam = float(z["lam"])
    N0, N1 = plan.N
    net = UNet(N1, N0, levels=int(z["levels"]), max_batch=1)
    net.set_params(z["theta"])
    dev = x.device
    b = torch.empty_like(x)
    Ab = torch.empty_like(x)
    xin = torch.empty((1, N1, N0), dtype=torch.complex64, device=dev)
    x0 = torch.empty((1, N1, N0), dtype=torch.complex64, device=dev)

    def plain(tol=tol):
        _, rep = plan.solve(y, maxit=args.maxit, tol=tol, lam=lam, out=x, toeplitz=True,
                            graph_iters=min(args.graph_iters, 8))
        return rep

    def warm(tol=tol):
        plan.EH(y, out=b)
        plan.normal(b, lam=lam, out=Ab, toeplitz=True)
        plan.sync()
        a0 = torch.vdot(b, b).real / torch.vdot(b, Ab).real
        xa = a0 * b
        s = torch.sqrt(torch.mean(xa.abs() ** 2))
        xin.copy_((xa / s).view(1, N1, N0))
        torch.cuda.synchronize()
        net.forward(xin, out=x0)
        x0.mul_(s)
        torch.cuda.synchronize()
        _, rep = plan.solve(y, x0=x0, maxit=args.maxit, tol=tol, lam=lam, out=x, toeplitz=True,
                            graph_iters=min(args.graph_iters, 8))
        return rep

    res = {"lambda": round(lam, 4), "network": os.path.relpath(args.weights, ROOT)}
    for name, fn in (("plain", plain), ("warm", warm)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = [fn() for _ in range(args.steps)]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        xr = x.cpu().numpy()
        res[name] = {"solves_per_s": round(args.steps / el, 3), "ms_per_solve": round(1e3 * el / args.steps, 3),
                     "iters": reps[-1]["max_iters"], "final_rel_resid": reps[-1]["final_rel"][0],
                     "rel_err_vs_phantom": round(float(np.linalg.norm(xr - xt) / np.linalg.norm(xt)), 5)}
    res["iteration_ratio"] = round(res["plain"]["iters"] / max(1, res["warm"]["iters"]), 3)
    res["speedup"] = round(res["warm"]["solves_per_s"] / res["plain"]["solves_per_s"], 3)
    # the same comparison at looser stopping tolerances (iterations only)
    res["iters_plain_warm_by_tol"] = {f"{t:g}": [plain(t)["max_iters"], warm(t)["max_iters"]]
                                      for t in (1e-3, 1e-4, tol)}
    res["note"] = ("workload item 0 is a training item; held_out_live are unseen seeds (items 5000+), "
                   "c5_protocol_500 the 500-seed evaluation of the training run")
    # held-out seeds, measured here (each with its own maps and noise)
    cur = torch.cuda.current_stream()
    try:
        from codeare_amd import warmstart_train as WT
        pipe = WT.Pipeline()   # makes its stream torch's current one
        rows = WT.evaluate(net, pipe, list(range(WT.EVAL_BASE, WT.EVAL_BASE + args.warm_seeds)), lam, tol,
                           log=lambda s: None)
        summ = WT.summarize(rows)
        res["held_out_live"] = {k: summ[k] for k in ("n", "iters_plain_mean", "iters_warm_mean", "iteration_ratio",
                                                     "solves_per_s_plain", "solves_per_s_warm", "r0_warm_median",
                                                     "r0_xa_median", "iters_saved_histogram")}
        res["held_out_live"]["speedup"] = round(summ["solves_per_s_warm"] / summ["solves_per_s_plain"], 3)
        del pipe
    except Exception as e:   # the headline numbers stand without it
        res["held_out_live"] = {"error": str(e)}
    finally:
        torch.cuda.synchronize()
        torch.cuda.set_stream(cur)
    rep_path = os.path.splitext(args.weights)[0] + "_report.json"
    if os.path.exists(rep_path):
        with open(rep_path) as f:
            rep = json.load(f)
        ev = rep["eval"]
        res["c5_protocol_500"] = {
            "train_items": rep["train_items"], "epochs": rep["epochs"], "loss": rep.get("loss", "l2"),
            "eval_items": ev["n"], "iters_plain_mean": ev["iters_plain_mean"], "iters_warm_mean": ev["iters_warm_mean"],
            "iters_plain_percentiles": ev.get("iters_plain_percentiles"),
            "iters_warm_percentiles": ev.get("iters_warm_percentiles"),
            "iters_saved_histogram": ev.get("iters_saved_histogram"),
            "r0_warm_median": ev.get("r0_warm_median"), "r0_xa_median": ev.get("r0_xa_median"),
            "warm_over_plain_time": round(ev["solves_per_s_warm"] / ev["solves_per_s_plain"], 3),
            "source": os.path.relpath(rep_path, ROOT)}
    return res


def dynamic_leg(args):
    """64 frames of 192^2 x 16 coils, 13..89 golden-angle spokes x 384 per frame,
    solved together (per-frame scalars and stopping) to 1e-5 with K-iteration
    HIP graphs; per-frame latency = host-observed time to the check that saw
    the frame stop."""
    import numpy as np
    import torch
    from codeare_amd import cgsense, data
    prob = data.make_problem(4, 0)
    wl = prob.wl
    M = prob.k.shape[1]
    t0 = time.perf_counter()
    plan = cgsense.Plan((wl.N, wl.N), M, wl.sigma, wl.m, ncoils=wl.coils, frame_offsets=prob.frame_offsets,
                        device_io=True, autotune=True)
    if torch.cuda.current_stream().cuda_stream == 0:
        torch.cuda.set_stream(torch.cuda.Stream())   # graphs are not captured on the null stream
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    plan.kspace(prob.k)
    plan.weights(prob.w)
    plan.sensitivities(prob.maps)
    t_plan = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    F = plan.F
    xt = torch.from_numpy(np.ascontiguousarray(prob.image.reshape(-1, order="F"))).to(dev)
    y = torch.empty(M * wl.coils, dtype=torch.complex64, device=dev)
    plan.E(xt, out=y)
    plan.sync()
    yh = data.add_noise(y.cpu().numpy(), wl.snr_db, prob.rng_state)
    y = torch.from_numpy(yh).to(dev)
    x = torch.empty(plan.Ntot * F, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    tol = wl.tol
    # lambda: the series has no stated value; lambda = 0 leaves the 13-spoke
    # frames singular (no convergence to 1e-5), so use 1e-3 of the largest
    # per-frame operator norm (power iteration over all frames at once)
    xs = torch.randn(plan.Ntot * F, dtype=torch.complex64, device=dev)
    qv = torch.empty_like(xs)
    for _ in range(20):
        plan.normal(xs, out=qv, toeplitz=True)
        plan.sync()
        nq = torch.linalg.vector_norm(qv.view(F, -1), dim=1)
        lmax_f = nq / torch.linalg.vector_norm(xs.view(F, -1), dim=1)
        xs = (qv.view(F, -1) / nq[:, None]).reshape(-1)
    lam = 1e-3 * float(lmax_f.max())

    def run():
        _, rep = plan.solve(y, maxit=args.maxit, tol=tol, lam=lam, out=x, toeplitz=True, graph_iters=8)
        return rep

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = run()
    el = time.perf_counter() - t0
    lat = np.array(rep["done_ms"][:F])
    its = np.array(rep["iters"][:F])
    counts = np.diff(prob.frame_offsets) // wl.samples
    return {"workload": wl.name, "frames": F, "spokes_per_frame": [int(counts.min()), int(counts.max())],
            "lambda": round(lam, 4), "lambda_rule": "1e-3 * max over frames of lambda_max(E^H W E)",
            "tol": tol, "graph_iters": 8,
            "solves_per_s": round(F / el, 3), "batch_ms": round(1e3 * el, 2),
            "frame_latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "frame_latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "iters_min_median_max": [int(its.min()), int(np.median(its)), int(its.max())],
            "final_rel_resid_max": float(np.max(rep["final_rel"][:F])),
            "frames_converged": int(np.sum(np.array(rep["final_rel"][:F]) <= tol)),
            "plan_setup_s": round(t_plan, 2)}


def volume_leg(args):
    """3D koosh ball: 128^3, 16 coils, 20000 x 256 samples, sigma 1.5, m 3,
    15 CG iterations from zero (lambda 0) through the Toeplitz operator."""
    import numpy as np
    import torch
    from codeare_amd import cgsense, data
    prob = data.make_problem(3, 0)
    wl = prob.wl
    M = prob.k.shape[1]
    t0 = time.perf_counter()
    plan = cgsense.Plan((wl.N,) * 3, M, wl.sigma, wl.m, ncoils=wl.coils, device_io=True, profile=True)
    if torch.cuda.current_stream().cuda_stream == 0:
        torch.cuda.set_stream(torch.cuda.Stream())   # graphs are not captured on the null stream
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    plan.kspace(prob.k)
    plan.weights(prob.w)
    plan.sensitivities(prob.maps)
    t_plan = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(np.ascontiguousarray(prob.image.reshape(-1, order="F"))).to(dev)
    y = torch.empty(M * wl.coils, dtype=torch.complex64, device=dev)
    plan.E(xt, out=y)
    x = torch.empty(plan.Ntot, dtype=torch.complex64, device=dev)
    res = {"workload": wl.name, "plan_setup_s": round(t_plan, 2)}
    for toep in (True, False):
        plan.solve(y, maxit=2, tol=0.0, lam=wl.lam, out=x, toeplitz=toep)
        plan.kernel_stats(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rep = plan.solve(y, maxit=wl.maxit, tol=0.0, lam=wl.lam, out=x, toeplitz=toep)
        el = time.perf_counter() - t0
        per = per_kernel(plan.kernel_stats(reset=True))
        res["toeplitz" if toep else "nufft"] = {
            "solve_s": round(el, 3), "iters": rep["max_iters"], "ms_per_iter": round(1e3 * el / wl.maxit, 2),
            "kernels": per}
    return res


def host_cpu():
    """(nproc, CPU model, threads the CPU legs use): the job's thread allowance
    (OMP_NUM_THREADS when set) or every core."""
    nproc = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or nproc
    except ValueError:
        threads = nproc
    return nproc, model, max(1, min(threads, nproc))


def _median_iter(marks, skip=1):
    """Median per-iteration time from the oracle's time marks, skipping the
    first `skip` iterations (warm-up)."""
    import numpy as np
    d = np.diff(marks)
    d = d[skip:] if d.size > skip else d
    return float(np.median(d)), int(d.size)


def cpu_baseline(prob, y, gpu_iters, lam, cpu_iters, ref_iters_gpu=None, ftiter=3, fteps=1e-3):
    """The oracle (the FP64-NFFT C port of the reference path: PRE_PSI-style
    precomputed window, FP32 CG vectors, OpenMP over coils) on the same inputs
    and the host's cores, on a bounded sample: E^H y plus `cpu_iters`
    iterations, each timed (steady clock); the first iteration is a warm-up and
    the solve time is E^H + (the GPU's iteration count) x the median iteration.
    Two modes, each against the GPU solve of the same mode: true adjoint (the
    headline's system) and codeare's own algorithm (reference E^H = ftiter
    steps of the NFFT3 CGNR per coil, CGLS with the squared-ratio stop,
    NCSENSE.hpp:251-301, CGLS.hpp:30-59)."""
    from oracle import ora
    ora.build()
    nproc, model, threads = host_cpu()
    ora.set_threads(threads)
    wl = prob.wl
    O = ora.NFFT((wl.N,) * wl.d, prob.k.shape[1], wl.sigma, wl.m)
    t0 = time.perf_counter()
    O.kspace(prob.k)
    t_setup = time.perf_counter() - t0
    O.cg_true(prob.maps, prob.w, y, maxit=1, tol=0.0, lam=lam)          # warm-up
    O.cg_true(prob.maps, prob.w, y, maxit=cpu_iters + 1, tol=0.0, lam=lam)
    mk = ora.time_marks()
    t_b = float(mk[0])
    t_it, n_it = _median_iter(mk)
    t_solve = t_b + gpu_iters * t_it
    out = {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": threads, "kind": "port",
           "mode": "true adjoint (the headline's system, through E then E^H)",
           "nproc": nproc, "cpu_model": model,
           "sample": f"E^H W y ({t_b:.3f} s) + {n_it + 1} CG iterations of the same 256^2 x 32-coil workload, "
                     f"each timed; median iteration {t_it:.3f} s (first one dropped as warm-up) x the "
                     f"{gpu_iters} iterations the GPU solve needed; trajectory setup {t_setup:.2f} s excluded",
           "eh_s": round(t_b, 4), "s_per_iter_median": round(t_it, 4)}
    if ref_iters_gpu:
        O.cgls_ref(prob.maps, prob.w, y, mode=1, ftiter=ftiter, fteps=fteps, maxit=1, cgeps=0.0, lam=lam)
        nref = max(3, cpu_iters // 2)
        O.cgls_ref(prob.maps, prob.w, y, mode=1, ftiter=ftiter, fteps=fteps, maxit=nref + 1, cgeps=0.0, lam=lam)
        mk = ora.time_marks()
        tb_r = float(mk[0])
        ti_r, n_r = _median_iter(mk)
        out["reference_algorithm"] = {
            "eh_s": round(tb_r, 4), "s_per_iter_median": round(ti_r, 4), "iters_timed": n_r,
            "solves_per_s": round(1.0 / (tb_r + ref_iters_gpu * ti_r), 6),
            "scaled_to_iters": ref_iters_gpu}
    return out


def fp64_errors(prob, y_noisy, x_gpu, iters, lam, lam_s=None, x_gpu_s=None, iters_s=None):
    """Error of the GPU images against the FP64 direct-NUDFT CG
    (oracle/nudft_torch.py) at the GPU's own iteration counts: the headline
    solve, the scaled-lambda solve and workload 1 (128^2, 8 coils, 10 CG
    iterations, lambda 0); for the headline also the part its data determine
    (DTFT inside the sampled disk |k| <= 0.5 on the 2N grid)."""
    import numpy as np
    import torch
    from codeare_amd import cgsense, data
    from oracle import nudft_torch as T
    dev = torch.device("cuda:0")
    wl = prob.wl
    N = wl.N
    S = torch.from_numpy(np.ascontiguousarray(prob.maps.astype(np.complex128))).to(dev)
    b = T.eh(prob.k, prob.w, y_noisy, S)
    Kh = T.kernel_spectrum(prob.k, prob.w, N, N, dev)

    def rel(a, b_):
        a = torch.as_tensor(a).to(dev).to(torch.complex128).reshape(-1)
        b_ = b_.reshape(-1)
        return float(torch.linalg.vector_norm(a - b_) / torch.linalg.vector_norm(b_))

    def disk(x):
        X = torch.fft.fft2(x, s=(2 * N, 2 * N))
        f = torch.fft.fftfreq(2 * N, device=dev, dtype=torch.float64)
        return X[(f[:, None] ** 2 + f[None, :] ** 2) <= 0.25]

    xg = torch.from_numpy(np.ascontiguousarray(x_gpu.reshape((N, N), order="F").astype(np.complex128))).to(dev)
    xo, n, _ = T.cg(S, Kh, b, lam, maxit=iters, tol=0.0)
    bn = float(torch.linalg.vector_norm(b))
    out = {"headline": {"iters": n, "rel_l2": round(rel(xg, xo), 6),
                        "rel_l2_inside_sampled_disk": round(float(torch.linalg.vector_norm(disk(xg) - disk(xo)) /
                                                                  torch.linalg.vector_norm(disk(xo))), 8),
                        "true_rel_resid_fp64": float(torch.linalg.vector_norm(b - T.normal(xg, S, Kh, lam))) / bn}}
    if x_gpu_s is not None:
        xgs = torch.from_numpy(np.ascontiguousarray(x_gpu_s.reshape((N, N), order="F").astype(np.complex128))).to(dev)
        xos, ns, _ = T.cg(S, Kh, b, lam_s, maxit=iters_s, tol=0.0)
        out["scaled_lambda"] = {"iters": ns, "rel_l2": round(rel(xgs, xos), 8)}
    # workload 1
    p1 = data.make_problem(1, 0)
    w1 = p1.wl
    P1 = cgsense.Plan((w1.N, w1.N), p1.k.shape[1], w1.sigma, w1.m, ncoils=w1.coils)
    P1.kspace(p1.k)
    P1.weights(p1.w)
    P1.sensitivities(p1.maps)
    y1 = data.make_data(p1, P1.E)
    x1, _ = P1.solve(y1, maxit=w1.maxit, tol=0.0, lam=w1.lam)
    S1 = torch.from_numpy(np.ascontiguousarray(p1.maps.astype(np.complex128))).to(dev)
    b1 = T.eh(p1.k, p1.w, y1, S1)
    K1 = T.kernel_spectrum(p1.k, p1.w, w1.N, w1.N, dev)
    xo1, n1, _ = T.cg(S1, K1, b1, w1.lam, maxit=w1.maxit, tol=0.0)
    out["workload1_cg10"] = {"iters": n1, "rel_l2": round(rel(np.ascontiguousarray(x1), xo1), 8)}
    return out


if __name__ == "__main__":
    main()
