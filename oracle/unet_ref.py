"""TEST INFRASTRUCTURE -- CPU restatement of the warm-start U-Net (forward and
backward) used only to check the HIP kernels in libcgsense_hip.so.

No counterpart in codeare (its CGLS, src/optimisation/CGLS.hpp:30-59, starts
from zero).  The network is the one the library implements:

  level l (l = 0..L-1, C_l = 32 * 2^l channels, side H / 2^l):
    a_l = relu(conv3x3(in_l)),  s_l = relu(conv3x3(a_l)),  in_{l+1} = maxpool2(s_l)
    in_0 = the complex64 input as channels (re, im), zero-padded to 16
  decoder l = L-2..0 (o_{L-1} = s_{L-1}):
    u_l = relu(conv3x3(concat(s_l, up2(o_{l+1})))),  o_l = relu(conv3x3(u_l))
  head: y = x + b + W o_0 (1x1, 32 -> 2, FP32)
  loss = mean over all real components of (y - t)^2

Parameters are one FP32 vector in the library's order (per conv: W[tap][co][ci]
with tap = 3*ky + kx, then bias; encoder levels, decoder levels L-2..0, head W
[2][32], head bias[2]).

With bf16=True the BF16 storage points of the GPU path are reproduced (conv
weights, every stored activation and activation gradient rounded to BF16,
sums in FP64); with bf16=False everything is FP64, which is what the
finite-difference check of the backward pass uses.
"""
from __future__ import annotations

import numpy as np

C0 = 32


def bf16_round(a):
    """Round float array to the nearest BF16 value (ties to even), as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_table(levels: int):
    """[(ca, cb, cout)] of the conv layers in parameter order."""
    t = []
    for l in range(levels):
        C = C0 << l
        cin = 16 if l == 0 else C0 << (l - 1)
        t += [(cin, 0, C), (C, 0, C)]
    for l in range(levels - 2, -1, -1):
        C = C0 << l
        t += [(C, 2 * C, C), (C, 0, C)]
    return t


def num_params(levels: int) -> int:
    n = sum(9 * co * (ca + cb) + co for ca, cb, co in layer_table(levels))
    return n + 2 * C0 + 2


def unpack(theta, levels):
    """Parameter vector -> list of (W[9][co][ci], b[co]) and head (W[2][32], b[2])."""
    theta = np.asarray(theta, dtype=np.float64)
    convs, off = [], 0
    for ca, cb, co in layer_table(levels):
        ci = ca + cb
        W = theta[off:off + 9 * co * ci].reshape(9, co, ci)
        off += 9 * co * ci
        b = theta[off:off + co]
        off += co
        convs.append((W, b))
    hw = theta[off:off + 2 * C0].reshape(2, C0)
    off += 2 * C0
    hb = theta[off:off + 2]
    return convs, (hw, hb)


def _shift(X, dy, dx):
    """X (B, H, W, C) -> X[:, y+dy-1, x+dx-1, :] with zero padding."""
    B, H, W, C = X.shape
    P = np.zeros((B, H + 2, W + 2, C), dtype=X.dtype)
    P[:, 1:-1, 1:-1] = X
    return P[:, dy:dy + H, dx:dx + W]


def conv3x3(X, W, b=None):
    """out[p, co] = b[co] + sum_tap sum_ci X[p + off(tap), ci] W[tap, co, ci]."""
    B, H, Wd, C = X.shape
    out = np.zeros((B, H, Wd, W.shape[1]))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        out += _shift(X, dy, dx).reshape(-1, C).dot(W[tap].T).reshape(B, H, Wd, -1)
    if b is not None:
        out += b
    return out


def conv3x3_wgrad(X, dY):
    """dW[tap, co, ci] = sum_p dY[p, co] X[p + off(tap), ci]."""
    C = X.shape[-1]
    dYf = dY.reshape(-1, dY.shape[-1])
    return np.stack([dYf.T.dot(_shift(X, *divmod(tap, 3)).reshape(-1, C)) for tap in range(9)])


def conv3x3_dgrad(dY, W):
    """Gradient w.r.t. X of conv3x3(X, W): the same convolution with the
    flipped, transposed kernel W'[8 - tap][ci][co]."""
    Wf = W[::-1].transpose(0, 2, 1)
    return conv3x3(dY, Wf)


def maxpool(X):
    B, H, W, C = X.shape
    return X.reshape(B, H // 2, 2, W // 2, 2, C).max(axis=(2, 4))


def maxpool_route(X, G):
    """Scatter G (pooled grid) to the first maximum of each 2x2 window of X
    (window order (0,0), (0,1), (1,0), (1,1)); windows whose maximum is <= 0
    get nothing (X is a ReLU output)."""
    B, H, W, C = X.shape
    out = np.zeros_like(X)
    win = [X[:, i::2, j::2] for i in (0, 1) for j in (0, 1)]
    best = win[0].copy()
    am = np.zeros(best.shape, dtype=np.int64)
    for q in range(1, 4):
        upd = win[q] > best
        best = np.where(upd, win[q], best)
        am = np.where(upd, q, am)
    live = best > 0
    for q, (i, j) in enumerate([(0, 0), (0, 1), (1, 0), (1, 1)]):
        out[:, i::2, j::2] = np.where((am == q) & live, G, 0.0)
    return out


def up2(X):
    return X.repeat(2, axis=1).repeat(2, axis=2)


def down_sum(G):
    B, H, W, C = G.shape
    return G.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


class UNetRef:
    def __init__(self, levels: int, theta, bf16: bool = True):
        self.L = levels
        self.bf = bf16
        self.convs, self.head = unpack(theta, levels)

    def _r(self, a):
        return bf16_round(a) if self.bf else a

    def _Wc(self, k):
        W, b = self.convs[k]
        return (bf16_round(W) if self.bf else W), b

    def forward(self, x, t=None):
        """x, t: complex (B, H, W).  Returns (y complex, loss or None)."""
        L, r = self.L, self._r
        B, H, W = x.shape
        xin = np.zeros((B, H, W, 16))
        xin[..., 0] = x.real
        xin[..., 1] = x.imag
        xin = r(xin)
        st = {"xin": xin}
        inp = xin
        for l in range(L):
            Wa, ba = self._Wc(2 * l)
            Wb, bb = self._Wc(2 * l + 1)
            a = r(np.maximum(conv3x3(inp, Wa, ba), 0))
            s = r(np.maximum(conv3x3(a, Wb, bb), 0))
            st[f"a{l}"], st[f"s{l}"] = a, s
            if l < L - 1:
                inp = maxpool(s)
                st[f"p{l}"] = inp
        below = st[f"s{L - 1}"]
        for l in range(L - 2, -1, -1):
            k = 2 * L + 2 * (L - 2 - l)
            Wa, ba = self._Wc(k)
            Wb, bb = self._Wc(k + 1)
            X = np.concatenate([st[f"s{l}"], up2(below)], axis=-1)
            u = r(np.maximum(conv3x3(X, Wa, ba), 0))
            o = r(np.maximum(conv3x3(u, Wb, bb), 0))
            st[f"u{l}"], st[f"o{l}"] = u, o
            below = o
        hw, hb = self.head
        f = st["o0"]
        yr = x.real + hb[0] + f.dot(hw[0])
        yi = x.imag + hb[1] + f.dot(hw[1])
        y = yr + 1j * yi
        self.st = st
        self.x = x
        loss = None
        if t is not None:
            n = 2 * y.size
            self.dout = (2.0 / n) * np.stack([yr - t.real, yi - t.imag], axis=-1)
            loss = float(np.sum((yr - t.real) ** 2 + (yi - t.imag) ** 2) / n)
        return y, loss

    def backward(self):
        """Gradient vector (parameter order) after forward(x, t)."""
        L, r, st = self.L, self._r, self.st
        grads = [None] * len(self.convs)
        hw, _ = self.head
        f = st["o0"]
        dout = self.dout
        g_head_w = np.stack([np.tensordot(dout[..., o], f, axes=([0, 1, 2], [0, 1, 2])) for o in range(2)])
        g_head_b = dout.sum(axis=(0, 1, 2))
        g_o = {0: r((f > 0) * (dout[..., :1] * hw[0] + dout[..., 1:] * hw[1]))}
        g_s = {}
        for l in range(0, L - 1):
            k = 2 * L + 2 * (L - 2 - l)
            Wa, _ = self._Wc(k)
            Wb, _ = self._Wc(k + 1)
            below = st[f"s{L - 1}"] if l + 1 == L - 1 else st[f"o{l + 1}"]
            u = st[f"u{l}"]
            grads[k + 1] = (conv3x3_wgrad(u, g_o[l]), g_o[l].sum(axis=(0, 1, 2)))
            g_u = r((u > 0) * conv3x3_dgrad(g_o[l], Wb))
            X = np.concatenate([st[f"s{l}"], up2(below)], axis=-1)
            grads[k] = (conv3x3_wgrad(X, g_u), g_u.sum(axis=(0, 1, 2)))
            dX = conv3x3_dgrad(g_u, Wa)
            C = st[f"s{l}"].shape[-1]
            g_s[l] = r((st[f"s{l}"] > 0) * dX[..., :C])
            g_up = r(dX[..., C:])
            g_b = r((below > 0) * down_sum(g_up))
            if l + 1 == L - 1:
                g_s[L - 1] = g_b
            else:
                g_o[l + 1] = g_b
        for l in range(L - 1, -1, -1):
            Wa, _ = self._Wc(2 * l)
            Wb, _ = self._Wc(2 * l + 1)
            a = st[f"a{l}"]
            inp = st["xin"] if l == 0 else st[f"p{l - 1}"]
            grads[2 * l + 1] = (conv3x3_wgrad(a, g_s[l]), g_s[l].sum(axis=(0, 1, 2)))
            g_a = r((a > 0) * conv3x3_dgrad(g_s[l], Wb))
            grads[2 * l] = (conv3x3_wgrad(inp, g_a), g_a.sum(axis=(0, 1, 2)))
            if l > 0:
                g_p = r(conv3x3_dgrad(g_a, Wa))
                g_s[l - 1] = r(g_s[l - 1] + maxpool_route(st[f"s{l - 1}"], g_p))
        out = []
        for gW, gb in grads:
            out += [gW.ravel(), gb.ravel()]
        out += [g_head_w.ravel(), g_head_b.ravel()]
        return np.concatenate(out)


def init_params(levels: int, seed: int) -> np.ndarray:
    """He-normal weights, zero biases, small random head (for tests)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ca, cb, co in layer_table(levels):
        ci = ca + cb
        fan = 9 * (2 if ca == 16 else ci)
        parts += [rng.normal(0, np.sqrt(2.0 / fan), 9 * co * ci), rng.normal(0, 0.05, co)]
    parts += [rng.normal(0, 0.1, 2 * C0), rng.normal(0, 0.05, 2)]
    return np.concatenate(parts).astype(np.float32)
