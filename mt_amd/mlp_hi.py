"""MLP-HI workflow on the engine: TEA-generated training data, device
training, deployment, and the diffusion-coefficient comparison that judges
it against plain TEA.  All arithmetic runs in the engine's HIP kernels
(include/mt_engine.h, MLP-HI section); this module only sequences calls.
"""
from __future__ import annotations

import time

import numpy as np

from .engine import Engine, HI_DEFAULT, HI_MLP, MODE_STREAMED
from . import params as P


def init_weights(seed: int = 0) -> np.ndarray:
    """Scaled normal initial weights (1/sqrt(fan-in)), zero biases."""
    rng = np.random.default_rng(seed)
    shapes = [(96, 128), (128,), (128, 128), (128,), (128, 128), (128,), (128, 9), (9,)]
    parts = [rng.normal(0, 1 / np.sqrt(s[0]), s).ravel() if len(s) == 2 else np.zeros(s) for s in shapes]
    return np.concatenate(parts).astype(np.float32)


def training_data(params, batch, frames: int, every: int, max_lj: int = 48):
    """Run TEA dynamics and collect (features, exact labels) of mobile beads
    every `every` steps; returns float32 arrays and the engine seconds."""
    eng = Engine(params, batch, max_lj=max_lj, mode=MODE_STREAMED)
    mob = (batch.fixed == 0) & (batch.extra == 0)
    X, Y = [], []
    t0 = time.perf_counter()
    for _ in range(frames):
        eng.run(every)
        f, t = eng.mlp_local()
        X.append(f[mob])
        Y.append(t[mob])
    dt = time.perf_counter() - t0
    eng.close()
    return np.concatenate(X), np.concatenate(Y), dt


def train(params, batch, X, Y, iters: int, bsize: int = 4096, lr: float = 1e-3, seed: int = 0):
    """Adam on the device (learning rate lr, then lr/3 after half the
    iterations and lr/10 for the last fifth); returns trained weights and the
    loss curve."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    X, Y = X[perm], Y[perm]
    eng = Engine(params, batch, mode=MODE_STREAMED)
    eng.mlp_set_weights(init_weights(seed))
    losses = []
    done = 0
    while done < iters:
        k = min(500, iters - done)
        scale = 1.0 if done < 0.5 * iters else (0.3 if done < 0.8 * iters else 0.1)
        losses.append(eng.mlp_train(X, Y, iters=k, batch=bsize, lr=lr * scale))
        done += k
    w = eng.mlp_get_weights()
    eng.close()
    return w, np.concatenate(losses)


def com_diffusion(batch, pos_start, pos_end, lag_ps):
    """Free-dimer centre-of-mass diffusion coefficient (nm^2/ps) from two
    unwrapped snapshots lag_ps apart: <|dR|^2> / (6 lag)."""
    a = batch.free_pairs
    c0 = 0.5 * (pos_start[a, :3].astype(np.float64) + pos_start[a + 1, :3])
    c1 = 0.5 * (pos_end[a, :3].astype(np.float64) + pos_end[a + 1, :3])
    d2 = ((c1 - c0) ** 2).sum(1)
    return float(d2.mean() / (6 * lag_ps)), float(d2.std() / np.sqrt(len(d2)) / (6 * lag_ps))


def run_mode(params, batch, mode: str, steps: int, warmup: int, weights=None, max_lj: int = 48, seed: int = 1234567,
             check_every: int = 100, tolerance: float = 0.02):
    """Run `steps` steps with 'tea', 'mlp' or 'none' hydrodynamics from the
    same seeds; returns (wall seconds, start and end states, mlp status)."""
    if mode == "none":
        params = P.derive({**_cfg_of(params), "tea_on": False})
    eng = Engine(params, batch, max_lj=max_lj, mode=MODE_STREAMED, rseed=seed)
    if mode == "mlp":
        eng.mlp_set_weights(weights)
        eng.set_hi_mode(HI_MLP, check_every=check_every, tolerance=tolerance)
    else:
        eng.set_hi_mode(HI_DEFAULT)
    eng.run(warmup)
    s0 = eng.get_state()
    t0 = time.perf_counter()
    eng.run(steps)
    dt = time.perf_counter() - t0
    s1 = eng.get_state()
    status = eng.mlp_status() if mode == "mlp" else None
    eng.close()
    return dt, s0, s1, status


def _cfg_of(params):
    """Config dict that reproduces `params` through params.derive (the
    free-draining variant keeps gammaR = 6 pi eta a with a = r_mon = tea_a)."""
    d = params.as_dict()
    return dict(P.BASE, dt=d["dt"], Temp=d["temp"], viscosity=d["gammaR"] / (6 * np.pi * 2.0), tea_a=d["tea_a"],
                tea_epsilon_freq=d["tea_epsilon_freq"], LJPairsCutoff=d["ljpairscutoff"],
                LJPairsUpdateFreq=d["ljpairsupdatefreq"])
