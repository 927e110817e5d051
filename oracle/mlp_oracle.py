"""FP64 numpy restatement of the MLP-HI pieces (no reference counterpart).

TEST INFRASTRUCTURE ONLY: the checker of the engine's MLP kernels.

  features  32 nearest non-extra beads of the same trajectory within
            r_c = 12 nm (minimum image), nearest first, relative position
            (r_j - r_i) / r_c; missing neighbours are zeros (96 values)
  target    A_i - I = sum_{j != i, r_ij < r_c} phi(r_ij) D_hat_ij, with the
            RPY block D_hat of src/bdhitea_kernel.cu:38-58 and the taper
            phi = 1 (r <= 8), (1 + cos(pi (r - 8) / 4)) / 2 (8 < r < 12)
  network   96 -> 128 -> 128 -> 128 -> 9, tanh hidden layers, linear out;
            weights flat: W1 (96x128, row k = input), b1, W2, b2, W3, b3,
            W4 (128x9), b4
  loss      mean over samples and outputs of (y - t)^2
"""
from __future__ import annotations

import numpy as np

NN, IN, H, OUT = 32, 96, 128, 9
RC, RIN = 12.0, 8.0
SHAPES = [(IN, H), (H,), (H, H), (H,), (H, H), (H,), (H, OUT), (OUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(w):
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(w[k:k + n], np.float64).reshape(s))
        k += n
    return out


def init_weights(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            parts.append(rng.normal(0, 1.0 / np.sqrt(s[0]), s).ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def round_bf16(w):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    a = np.asarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward(w, x, keep=False):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    x = np.asarray(x, np.float64)
    h1 = np.tanh(x @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    h3 = np.tanh(h2 @ W3 + b3)
    y = h3 @ W4 + b4
    return (y, (x, h1, h2, h3)) if keep else y


def loss_and_grad(w, x, t):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    y, (x, h1, h2, h3) = forward(w, x, keep=True)
    n = y.size
    d = y - np.asarray(t, np.float64)
    loss = float((d * d).sum() / n)
    g4 = 2 * d / n
    gW4, gb4 = h3.T @ g4, g4.sum(0)
    g3 = (g4 @ W4.T) * (1 - h3 * h3)
    gW3, gb3 = h2.T @ g3, g3.sum(0)
    g2 = (g3 @ W3.T) * (1 - h2 * h2)
    gW2, gb2 = h1.T @ g2, g2.sum(0)
    g1 = (g2 @ W2.T) * (1 - h1 * h1)
    gW1, gb1 = x.T @ g1, g1.sum(0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, grad


def taper(r):
    r = np.asarray(r, np.float64)
    return np.where(r <= RIN, 1.0, np.where(r < RC, 0.5 * (1 + np.cos(np.pi * (r - RIN) / (RC - RIN))), 0.0))


def rpy_block(d, a):
    r = np.linalg.norm(d)
    e = d / r
    ra = r / a
    if ra > 2:
        cii, crr = 0.75 / ra * (1 + 2 / (3 * ra * ra)), 0.75 / ra * (1 - 2 / (ra * ra))
    else:
        cii, crr = 1 - 9 * ra / 32, 3 * ra / 32
    return cii * np.eye(3) + crr * np.outer(e, e)


def local_features_targets(coords, off, extra, box, a):
    """Brute-force features (nb, 96) and targets (nb, 9) in FP64."""
    pos = np.asarray(coords, np.float64)[:, :3]
    nb = len(pos)
    feat = np.zeros((nb, IN))
    targ = np.zeros((nb, OUT))
    for t in range(len(off) - 1):
        s, e = int(off[t]), int(off[t + 1])
        L = float(box[t]) if box is not None else 0.0
        for i in range(s, e):
            if extra[i]:
                continue
            d = pos[s:e] - pos[i]
            if L > 0:
                d -= L * np.rint(d / L)
            r = np.linalg.norm(d, axis=1)
            ok = (r < RC) & (np.arange(s, e) != i) & (np.asarray(extra[s:e]) == 0)
            idx = np.nonzero(ok)[0]
            order = idx[np.lexsort((idx, r[idx]))][:NN]
            for k, j in enumerate(order):
                feat[i, 3 * k:3 * k + 3] = d[j] / RC
            A = np.zeros((3, 3))
            for j in idx:
                A += taper(r[j]) * rpy_block(d[j], a)
            targ[i] = A.ravel()
    return feat, targ
