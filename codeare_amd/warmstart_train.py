"""Train and evaluate the CG warm-start U-Net on the 256^2, 32-coil golden-angle
radial workload (402 spokes x 512 samples, sigma 2, m 3, lambda 1e-3).

Training pairs come from the GPU solver itself: for each seeded phantom
(make_problem(5, item), coil maps and noise from the same seed) the input is
x_a = alpha0 E^H W y and the target is the CG image converged to relative
residual `target_tol` (norm ratio).  Both are divided by RMS(x_a).  The
network is trained with Adam (lr 1e-3) on mini-batches by the library's own
forward/backward kernels, then evaluated on held-out items (item >= 5000):
CG from the network output vs CG from zero, both to `tol` with the same test.

    python -m codeare_amd.warmstart_train --train 512 --eval 32 --epochs 20 \
        --out codeare_amd/weights/unet_golden2d_256_c32.npz
"""
from __future__ import annotations

import argparse
import json
import queue
import threading
import time

import numpy as np

from . import data
from .cgsense import Plan
from .warmstart import UNet, batch_to_image, images_to_batch, network_input

LAM = 1e-3
EVAL_BASE = 5000


def make_plan(cfg: int = 5) -> Plan:
    prob = data.make_problem(cfg, 0)
    wl = prob.wl
    P = Plan((wl.N, wl.N), prob.k.shape[1], wl.sigma, wl.m, ncoils=wl.coils)
    P.kspace(prob.k)
    P.weights(prob.w)
    return P


def _problems(items, cfg=5):
    """Host-side generation of phantoms/maps in a helper thread (overlaps the GPU)."""
    q: queue.Queue = queue.Queue(maxsize=4)

    def work():
        for it in items:
            q.put((it, data.make_problem(cfg, it)))
        q.put(None)

    threading.Thread(target=work, daemon=True).start()
    while True:
        v = q.get()
        if v is None:
            return
        yield v


def generate_pairs(P: Plan, items, target_tol=1e-7, maxit=20000, graph_iters=16, log=print):
    X, T, its = [], [], []
    t0 = time.perf_counter()
    for n, (it, prob) in enumerate(_problems(items)):
        P.sensitivities(prob.maps)
        y = data.make_data(prob, P.E)
        xa, s, _ = network_input(P, y, LAM)
        xt, rep = P.solve(y, maxit=maxit, tol=target_tol, lam=LAM, toeplitz=True, graph_iters=graph_iters)
        X.append(xa / s)
        T.append(xt / s)
        its.append(rep["iters"][0])
        if (n + 1) % 32 == 0:
            log(f"  pairs {n + 1}/{len(items)}: {time.perf_counter() - t0:.1f} s, "
                f"CG iterations to {target_tol:g}: mean {np.mean(its):.0f}")
    return images_to_batch(X), images_to_batch(T), its


def train(net: UNet, X, T, epochs=20, batch=16, lr=1e-3, seed=0, log=print):
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(n)
        ls = []
        for s in range(0, n - batch + 1, batch):
            idx = order[s:s + batch]
            ls.append(net.train_step(X[idx], T[idx], lr=lr))
        hist.append(float(np.mean(ls)))
        log(f"  epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}  ({time.perf_counter() - t0:.1f} s)")
    return hist


def evaluate(net: UNet, P: Plan, items, tol=1e-5, maxit=20000, graph_iters=16, log=print):
    rows = []
    for it, prob in _problems(items):
        P.sensitivities(prob.maps)
        y = data.make_data(prob, P.E)
        P.solve(y, maxit=8, tol=tol, lam=LAM, toeplitz=True, graph_iters=graph_iters)  # warm caches
        t0 = time.perf_counter()
        xp, rp = P.solve(y, maxit=maxit, tol=tol, lam=LAM, toeplitz=True, graph_iters=graph_iters)
        tp = time.perf_counter() - t0
        t0 = time.perf_counter()
        xa, s, _ = network_input(P, y, LAM)
        out = net.forward(images_to_batch([xa / s]))
        x0 = (batch_to_image(out[0]) * np.float32(s)).astype(np.complex64)
        tn = time.perf_counter() - t0
        t0 = time.perf_counter()
        xw, rw = P.solve(y, x0=x0, maxit=maxit, tol=tol, lam=LAM, toeplitz=True, graph_iters=graph_iters)
        tw = time.perf_counter() - t0
        ref = prob.image
        rows.append({
            "item": it, "iters_plain": rp["iters"][0], "iters_warm": rw["iters"][0],
            "s_plain": tp, "s_warm": tw + tn, "s_net": tn,
            "rel_plain": rp["final_rel"][0], "rel_warm": rw["final_rel"][0],
            "warm_vs_plain": float(np.linalg.norm(xw - xp) / np.linalg.norm(xp)),
            "x0_vs_plain": float(np.linalg.norm(x0 - xp) / np.linalg.norm(xp)),
            "xa_vs_plain": float(np.linalg.norm(xa - xp) / np.linalg.norm(xp)),
            "plain_vs_phantom": float(np.linalg.norm(xp - ref) / np.linalg.norm(ref)),
        })
        r = rows[-1]
        log(f"  eval item {it}: iters plain {r['iters_plain']} warm {r['iters_warm']}; "
            f"x0 err {r['x0_vs_plain']:.3e} (x_a {r['xa_vs_plain']:.3e})")
    return rows


def summarize(rows):
    ip = np.array([r["iters_plain"] for r in rows], float)
    iw = np.array([r["iters_warm"] for r in rows], float)
    sp = np.array([r["s_plain"] for r in rows])
    sw = np.array([r["s_warm"] for r in rows])
    return {
        "n": len(rows),
        "iters_plain_mean": float(ip.mean()), "iters_warm_mean": float(iw.mean()),
        "iteration_ratio": float(ip.mean() / iw.mean()),
        "solves_per_s_plain": float(len(rows) / sp.sum()), "solves_per_s_warm": float(len(rows) / sw.sum()),
        "x0_err_mean": float(np.mean([r["x0_vs_plain"] for r in rows])),
        "xa_err_mean": float(np.mean([r["xa_vs_plain"] for r in rows])),
        "warm_vs_plain_max": float(np.max([r["warm_vs_plain"] for r in rows])),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--eval", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target-tol", type=float, default=1e-7)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--out", default="")
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)

    P = make_plan()
    net = UNet(256, 256, levels=4, max_batch=a.batch, seed=20261202)
    log(f"U-Net: {net.nparams} parameters")
    log(f"generating {a.train} training pairs (CG to {a.target_tol:g})")
    t0 = time.perf_counter()
    X, T, its = generate_pairs(P, list(range(a.train)), a.target_tol, log=log)
    t_gen = time.perf_counter() - t0
    log(f"training: {a.epochs} epochs, batch {a.batch}")
    net.profile(1)
    t0 = time.perf_counter()
    hist = train(net, X, T, a.epochs, a.batch, a.lr, log=log)
    t_train = time.perf_counter() - t0
    prof = net.profile(0)
    if a.out:
        net.save(a.out)
    log(f"evaluating on {a.eval} held-out items")
    rows = evaluate(net, P, list(range(EVAL_BASE, EVAL_BASE + a.eval)), a.tol, log=log)
    summ = summarize(rows)
    mfma = {k: (prof[k + "_flop"] / (prof[k + "_ms"] * 1e-3) / 1e12 if prof[k + "_ms"] > 0 else 0.0)
            for k in ("fwd", "dgrad", "wgrad")}
    rep = {"train_items": a.train, "eval_items": a.eval, "epochs": a.epochs, "batch": a.batch,
           "lr": a.lr, "target_tol": a.target_tol, "tol": a.tol, "lambda": LAM,
           "pair_generation_s": t_gen, "train_s": t_train, "target_cg_iters_mean": float(np.mean(its)),
           "loss_history": hist, "conv_TFLOPs": mfma, "kernel_ms": prof, "eval": summ, "rows": rows}
    log(json.dumps({"eval": summ, "conv_TFLOPs": mfma}))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
