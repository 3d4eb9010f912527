"""Train and evaluate the CG warm-start U-Net on the 256^2, 32-coil golden-angle
radial workload (402 spokes x 512 samples, sigma 2, m 3).

Regularisation.  With codeare's unnormalised NFFT and w = |k| the normal
operator has lambda_max ~ 1e5, so the literal lambda = 1e-3 leaves the
k-space corners outside the sampled disk unregularised: CG to 1e-5 fits the
noise there and "CG converged to 1e-7" does not exist in FP32 within 20000
iterations.  The learned warm start is therefore trained and judged on the
well-posed problem lambda = 1e-3 * lambda_max (power iteration on the first
item; stored with the weights so the benchmark uses the same value).

Training pairs come from the GPU solver itself: for each seeded phantom
(make_problem(5, item): ellipses and coil maps from the item's seed, noise at
40 dB drawn on the GPU from the same seed) the input is
x_a = alpha0 E^H W y and the target the CG image at relative residual
`target_tol`; both divided by RMS(x_a).  Adam (lr 1e-3) on mini-batches of 16
with the library's own forward/backward kernels.  Evaluation on held-out items
(item >= 5000): CG from the network output vs CG from zero, both to `tol`
with the same test, timed on the GPU.

    python -m codeare_amd.warmstart_train --train 1024 --eval 32 --epochs 20 \\
        --out codeare_amd/weights/unet_golden2d_256_c32.npz
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import math
import time

import numpy as np

from . import data
from .cgsense import Plan
from .warmstart import UNet

EVAL_BASE = 5000


class Pipeline:
    """Device-resident data path for one trajectory (the workload's)."""

    def __init__(self, cfg: int = 5):
        import torch
        self.torch = torch
        prob = data.make_problem(cfg, 0)
        self.cfg = cfg
        self.wl = prob.wl
        self.N = prob.wl.N
        self.M = prob.k.shape[1]
        self.C = prob.wl.coils
        self.P = Plan((self.N, self.N), self.M, prob.wl.sigma, prob.wl.m, ncoils=self.C, device_io=True)
        # one stream for torch's tensor work and the plan's kernels, so tensors
        # torch writes are ready when the plan reads them (and vice versa)
        self.stream = torch.cuda.Stream()
        torch.cuda.set_stream(self.stream)
        self.P.set_stream(self.stream.cuda_stream)
        self.P.kspace(prob.k)
        self.P.weights(prob.w)
        self.P.sensitivities(prob.maps)
        self.dev = torch.device("cuda:0")
        t = torch
        self.y = t.empty(self.M * self.C, dtype=t.complex64, device=self.dev)
        self.x = t.empty(self.N * self.N, dtype=t.complex64, device=self.dev)
        self.b = t.empty_like(self.x)
        self.Ab = t.empty_like(self.x)
        self.xs = t.empty_like(self.x)

    def lambda_max(self, iters=30):
        t = self.torch
        g = t.Generator(device=self.dev)
        g.manual_seed(1)
        v = t.randn(self.N * self.N, dtype=t.complex64, device=self.dev, generator=g)
        lm = 0.0
        for _ in range(iters):
            self.P.normal(v, out=self.Ab, toeplitz=True)
            self.P.sync()
            lm = float(t.linalg.vector_norm(self.Ab) / t.linalg.vector_norm(v))
            v = self.Ab / t.linalg.vector_norm(self.Ab)
        return lm

    def load(self, prob, item):
        """Maps of the item, y = E x + noise (GPU, seeded by the item)."""
        t = self.torch
        self.P.sensitivities(prob.maps)
        xt = t.from_numpy(np.ascontiguousarray(prob.image.reshape(-1, order="F"))).to(self.dev)
        self.P.E(xt, out=self.y)
        self.P.sync()
        rms = float(t.sqrt(t.mean(self.y.abs() ** 2)))
        sd = rms * 10 ** (-self.wl.snr_db / 20.0) / math.sqrt(2.0)
        g = t.Generator(device=self.dev)
        g.manual_seed(data.seed_for(self.cfg, item))
        self.y += (sd * t.randn(self.y.shape, dtype=t.complex64, device=self.dev, generator=g)).to(t.complex64)
        t.cuda.synchronize()
        return xt

    def net_input(self, lam):
        """x_a = alpha0 E^H W y on the device (in self.b), and RMS(x_a)."""
        t = self.torch
        self.P.EH(self.y, out=self.b)
        self.P.normal(self.b, lam=lam, out=self.Ab, toeplitz=True)
        self.P.sync()
        a0 = t.vdot(self.b, self.b).real / t.vdot(self.b, self.Ab).real
        self.b.mul_(a0)
        s = float(t.sqrt(t.mean(self.b.abs() ** 2)))
        return self.b, s

    def solve(self, lam, tol, maxit, x0=None, graph_iters=16):
        _, rep = self.P.solve(self.y, x0=x0, maxit=maxit, tol=tol, lam=lam, out=self.x, toeplitz=True,
                              graph_iters=graph_iters)
        return self.x, rep


def _problems(items, cfg=5, workers=8):
    """Phantoms and maps generated on host threads, in order, a few ahead."""
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        futs = {}
        nxt = 0
        for i, it in enumerate(items):
            while nxt < len(items) and nxt < i + 2 * workers:
                futs[nxt] = ex.submit(data.make_problem, cfg, items[nxt])
                nxt += 1
            yield it, futs.pop(i).result()


def _as_hw(v, N):
    """Device image (column-major N0 x N1, flat) -> host (H=N1, W=N0) array."""
    return v.view(N, N).cpu().numpy()


def generate_pairs(pipe: Pipeline, items, lam, target_tol=1e-7, maxit=2000, log=print, keep_maps=False):
    """(X, T, iterations, maps): network inputs and CG targets; with keep_maps
    the items' maps as one device tensor (n, Ntot * C) (16.8 MB per item at
    256^2 x 32 coils)."""
    X, T, its = [], [], []
    t0 = time.perf_counter()
    N = pipe.N
    maps = None
    if keep_maps:
        t = pipe.torch
        maps = t.empty((len(items), N * N * pipe.C), dtype=t.complex64, device=pipe.dev)
    for n, (it, prob) in enumerate(_problems(items)):
        if keep_maps:
            maps[n].copy_(pipe.torch.from_numpy(np.asarray(prob.maps).reshape(-1, order="F")))
        pipe.load(prob, it)
        xa, s = pipe.net_input(lam)
        X.append(_as_hw(xa, N) / s)
        xt, rep = pipe.solve(lam, target_tol, maxit)
        T.append(_as_hw(xt, N) / s)
        its.append(rep["iters"][0])
        if (n + 1) % 64 == 0:
            log(f"  pairs {n + 1}/{len(items)}: {time.perf_counter() - t0:.1f} s, "
                f"CG iterations to {target_tol:g}: mean {np.mean(its):.0f}")
    return np.stack(X).astype(np.complex64), np.stack(T).astype(np.complex64), its, maps


def train(net: UNet, X, T, epochs=20, batch=16, lr=1e-3, seed=0, log=print, loss="l2", pipe=None, lam=None,
          lmax=None, maps=None):
    """Adam on mini-batches.

    loss "l2": the image error |x0 - x*|^2 (the library's own training step).
    loss "anorm": the A-norm error e^H A e, e = x0 - x*, A = (E^H W E +
    lambda) / lambda_max with each item's own maps; its gradient 2 A e takes
    one Toeplitz normal-operator application per image.  loss "resid": the
    normal-equation residual |A e|^2 = |A x0 - A x*|^2, CG's stopping
    quantity; gradient 2 A^2 e, two applications (A is self-adjoint).  Both
    enter the library's L2 step through the shifted target t' = x0 - g / 2
    (g the loss gradient), whose L2 gradient is g up to a constant factor.
    X, T: (n, H, W) complex64 host arrays; maps: device tensor (n, Ntot * C)
    of the items' maps for the operator losses."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    hist = []
    t0 = time.perf_counter()
    if loss != "l2":
        torch = pipe.torch
        dev = pipe.dev
        Xd = torch.from_numpy(X).to(dev)
        Td = torch.from_numpy(T).to(dev)
        H, W = Xd.shape[1], Xd.shape[2]
        out = torch.empty((batch, H, W), dtype=torch.complex64, device=dev)
        tp = torch.empty_like(out)
        q = torch.empty(H * W, dtype=torch.complex64, device=dev)
        q2 = torch.empty_like(q)
    for ep in range(epochs):
        order = rng.permutation(n)
        ls = []
        for s in range(0, n - batch + 1, batch):
            idx = np.sort(order[s:s + batch])
            if loss == "l2":
                ls.append(net.train_step(X[idx], T[idx], lr=lr))
                continue
            ii = torch.from_numpy(idx).to(dev)
            xb = Xd.index_select(0, ii).contiguous()
            torch.cuda.synchronize()
            net.forward(xb, out=out)
            torch.cuda.synchronize()
            e = out - Td.index_select(0, ii)
            val = torch.zeros((), dtype=torch.float64, device=dev)
            for b in range(batch):
                eb = e[b].reshape(-1).contiguous()
                pipe.P.sensitivities_device(maps[int(idx[b])])
                pipe.P.normal(eb, lam=lam, out=q, toeplitz=True)
                if loss == "anorm":
                    q.div_(lmax)
                    val += torch.vdot(eb, q).real.double()
                else:
                    pipe.P.normal(q, lam=lam, out=q2, toeplitz=True)
                    val += (q.abs() ** 2).sum().double() / (lmax * lmax)
                    q.copy_(q2).div_(lmax * lmax)
                tp[b] = out[b] - q.view(H, W)
            torch.cuda.synchronize()
            net.train_step(xb, tp, lr=lr)
            ls.append(float(val) / (batch * H * W))
        hist.append(float(np.mean(ls)))
        log(f"  epoch {ep + 1}/{epochs}: loss ({loss}) {hist[-1]:.4e}  ({time.perf_counter() - t0:.1f} s)")
    return hist


def evaluate(net: UNet, pipe: Pipeline, items, lam, tol=1e-5, maxit=3000, log=print):
    t = pipe.torch
    N = pipe.N
    rows = []
    xin = t.empty((1, N, N), dtype=t.complex64, device=pipe.dev)
    x0 = t.empty_like(xin)
    b = t.empty(N * N, dtype=t.complex64, device=pipe.dev)
    q = t.empty_like(b)

    def r0(v):
        """||b - (A + lambda) v|| / ||b||: the relative residual CG starts from."""
        pipe.P.normal(v.reshape(-1).contiguous(), lam=lam, out=q, toeplitz=True)
        pipe.P.sync()
        return float(t.linalg.vector_norm(b - q) / t.linalg.vector_norm(b))

    for it, prob in _problems(items):
        xt = pipe.load(prob, it)
        pipe.solve(lam, tol, 16)                                   # warm caches
        t.cuda.synchronize()
        t0 = time.perf_counter()
        xp, rp = pipe.solve(lam, tol, maxit)
        t.cuda.synchronize()
        tp = time.perf_counter() - t0
        xp = xp.clone()
        t0 = time.perf_counter()
        xa, s = pipe.net_input(lam)
        xin.copy_((xa / s).view(1, N, N))
        t.cuda.synchronize()
        net.forward(xin, out=x0)
        x0.mul_(s)
        t.cuda.synchronize()
        tn = time.perf_counter() - t0
        t0 = time.perf_counter()
        xw, rw = pipe.solve(lam, tol, maxit, x0=x0)
        t.cuda.synchronize()
        tw = time.perf_counter() - t0
        pipe.P.EH(pipe.y, out=b)
        pipe.P.sync()
        r0_warm, r0_xa = r0(x0), r0(xa)

        def rel(a, b):
            return float(t.linalg.vector_norm(a.reshape(-1) - b.reshape(-1)) / t.linalg.vector_norm(b.reshape(-1)))

        rows.append({
            "item": it, "iters_plain": rp["iters"][0], "iters_warm": rw["iters"][0],
            "s_plain": tp, "s_warm": tw + tn, "s_net": tn,
            "rel_plain": rp["final_rel"][0], "rel_warm": rw["final_rel"][0],
            "warm_vs_plain": rel(xw, xp), "x0_vs_plain": rel(x0, xp), "xa_vs_plain": rel(xa, xp),
            "plain_vs_phantom": rel(xp, xt), "warm_vs_phantom": rel(xw, xt),
            "r0_warm": r0_warm, "r0_xa": r0_xa,
        })
        r = rows[-1]
        log(f"  eval item {it}: iters plain {r['iters_plain']} warm {r['iters_warm']}; "
            f"x0 err {r['x0_vs_plain']:.3e} (x_a {r['xa_vs_plain']:.3e}); warm vs plain {r['warm_vs_plain']:.2e}; "
            f"starting residual warm {r0_warm:.3e} (x_a {r0_xa:.3e}, zero 1)")
    return rows


def summarize(rows):
    ip = np.array([r["iters_plain"] for r in rows], float)
    iw = np.array([r["iters_warm"] for r in rows], float)
    sp = np.array([r["s_plain"] for r in rows])
    sw = np.array([r["s_warm"] for r in rows])
    q = [0, 5, 25, 50, 75, 95, 100]
    return {
        "n": len(rows),
        "iters_plain_mean": float(ip.mean()), "iters_warm_mean": float(iw.mean()),
        "iters_plain_percentiles": dict(zip(map(str, q), np.percentile(ip, q).tolist())),
        "iters_warm_percentiles": dict(zip(map(str, q), np.percentile(iw, q).tolist())),
        "iters_saved_histogram": {str(int(k)): int(v) for k, v in zip(*np.unique(ip - iw, return_counts=True))},
        "r0_warm_median": float(np.median([r["r0_warm"] for r in rows])) if "r0_warm" in rows[0] else None,
        "r0_xa_median": float(np.median([r["r0_xa"] for r in rows])) if "r0_xa" in rows[0] else None,
        "iteration_ratio": float(ip.mean() / iw.mean()),
        "solves_per_s_plain": float(len(rows) / sp.sum()), "solves_per_s_warm": float(len(rows) / sw.sum()),
        "x0_err_mean": float(np.mean([r["x0_vs_plain"] for r in rows])),
        "xa_err_mean": float(np.mean([r["xa_vs_plain"] for r in rows])),
        "warm_vs_plain_max": float(np.max([r["warm_vs_plain"] for r in rows])),
        "plain_vs_phantom_mean": float(np.mean([r["plain_vs_phantom"] for r in rows])),
        "warm_vs_phantom_mean": float(np.mean([r["warm_vs_phantom"] for r in rows])),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--train", type=int, default=1024)
    ap.add_argument("--eval", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target-tol", type=float, default=1e-7)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--lambda-rel", type=float, default=1e-3, help="lambda = this x lambda_max")
    ap.add_argument("--loss", choices=("l2", "anorm", "resid"), default="l2",
                    help="training loss: image L2 error, the A-norm error, or the normal-equation residual")
    ap.add_argument("--out", default="")
    ap.add_argument("--report", default="")
    a = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)

    pipe = Pipeline()
    lmax = pipe.lambda_max()
    lam = a.lambda_rel * lmax
    log(f"lambda_max {lmax:.1f}, lambda {lam:.3f}")
    net = UNet(pipe.N, pipe.N, levels=4, max_batch=a.batch, seed=20261202)
    log(f"U-Net: {net.nparams} parameters")
    log(f"generating {a.train} training pairs (CG to {a.target_tol:g})")
    t0 = time.perf_counter()
    X, T, its, maps = generate_pairs(pipe, list(range(a.train)), lam, a.target_tol, log=log,
                                     keep_maps=a.loss != "l2")
    t_gen = time.perf_counter() - t0
    log(f"training: {a.epochs} epochs, batch {a.batch}")
    net.profile(1)
    t0 = time.perf_counter()
    hist = train(net, X, T, a.epochs, a.batch, a.lr, log=log, loss=a.loss, pipe=pipe, lam=lam, lmax=lmax,
                 maps=maps)
    del maps
    t_train = time.perf_counter() - t0
    prof = net.profile(0)
    if a.out:
        np.savez(a.out, theta=net.params(), H=net.H, W=net.W, levels=net.levels, lam=lam, lambda_max=lmax,
                 loss=a.loss, train_items=a.train, epochs=a.epochs)
    log(f"evaluating on {a.eval} held-out items")
    rows = evaluate(net, pipe, list(range(EVAL_BASE, EVAL_BASE + a.eval)), lam, a.tol, log=log)
    summ = summarize(rows)
    mfma = {k: (prof[k + "_flop"] / (prof[k + "_ms"] * 1e-3) / 1e12 if prof[k + "_ms"] > 0 else 0.0)
            for k in ("fwd", "dgrad", "wgrad")}
    rep = {"train_items": a.train, "eval_items": a.eval, "epochs": a.epochs, "batch": a.batch, "loss": a.loss,
           "lr": a.lr, "target_tol": a.target_tol, "tol": a.tol, "lambda": lam, "lambda_max": lmax,
           "pair_generation_s": t_gen, "train_s": t_train, "target_cg_iters_mean": float(np.mean(its)),
           "loss_history": hist, "conv_TFLOPs": mfma, "kernel_ms": prof, "eval": summ, "rows": rows}
    log(json.dumps({"eval": summ, "conv_TFLOPs": mfma}))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
