"""FP64 Ermak-McCammon with the full Rotne-Prager-Yamakawa diffusion tensor.

TEST INFRASTRUCTURE ONLY (the checker TEA results are judged against).

The reference declares an "exact" (Cholesky) HI mode but never implements it
(src/mt.h:113,119; src/bdhitea.cuh:16), so this is written from the
mathematics:
  D = (kB T / gammaR) D_hat,  D_hat_ii = I,
  D_hat_ij = (3a/4r)[(1 + 2a^2/3r^2) I + (1 - 2a^2/r^2) e e^T]   (r >= 2a)
           = (1 - 9r/32a) I + (3r/32a) e e^T                      (r <  2a)
  dr = (dt / kB T) D F + sqrt(2 dt) L xi,  L L^T = D.
TEA (Geyer & Winter 2009) keeps the deterministic part exactly,
(dt/gammaR)(F_i + sum_j D_hat_ij F_j), and approximates L by
B_ii = C_i, B_ij = beta C_i D_hat_ij (src/bdhitea_kernel.cu:104-213).
"""
from __future__ import annotations

import numpy as np

KB = 0.0019872041


def rpy_hat(pos: np.ndarray, a: float) -> np.ndarray:
    """D_hat as a dense (3N, 3N) FP64 matrix."""
    pos = np.asarray(pos, np.float64)
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]            # d[i, j] = r_j - r_i
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, 1.0)
    e = d / r[:, :, None]
    ra = r / a
    far = ra > 2.0
    cii = np.where(far, 0.75 / ra * (1 + 2.0 / (3.0 * ra * ra)), 1 - 9 * ra / 32)
    crr = np.where(far, 0.75 / ra * (1 - 2.0 / (ra * ra)), 3 * ra / 32)
    blocks = cii[:, :, None, None] * np.eye(3) + crr[:, :, None, None] * e[:, :, :, None] * e[:, :, None, :]
    idx = np.arange(n)
    blocks[idx, idx] = np.eye(3)
    return blocks.transpose(0, 2, 1, 3).reshape(3 * n, 3 * n)


def drift(pos, F, a, dt, gammaR):
    """Deterministic EM displacement (dt/gammaR) D_hat F, FP64, (N, 3)."""
    Dh = rpy_hat(pos, a)
    return dt / gammaR * (Dh @ np.asarray(F, np.float64).reshape(-1)).reshape(-1, 3)


def em_step(pos, F, xi, a, dt, gammaR, temp):
    """One exact Cholesky EM step; xi: (N, 3) standard normals."""
    Dh = rpy_hat(pos, a)
    L = np.linalg.cholesky(Dh)
    det = dt / gammaR * (Dh @ np.asarray(F, np.float64).reshape(-1))
    sto = np.sqrt(2 * KB * temp * dt / gammaR) * (L @ np.asarray(xi, np.float64).reshape(-1))
    return np.asarray(pos, np.float64) + (det + sto).reshape(-1, 3)


def tea_b_matrix(pos, a):
    """TEA's noise matrix B (3N, 3N) with its FP64 epsilon and beta."""
    Dh = rpy_hat(pos, a)
    n3 = Dh.shape[0]
    off = Dh.copy()
    for i in range(0, n3, 3):
        off[i:i + 3, i:i + 3] = 0
    eps = off.sum() / (n3 * (n3 - 3))
    aa = (n3 - 1) * eps * eps - (n3 - 2) * eps
    beta = 0.5 if abs(aa) < 1e-7 else (1 - np.sqrt(1 - aa)) / aa
    ci = (off ** 2).sum(1)
    C = 1.0 / np.sqrt(1 + beta * beta * ci)
    B = beta * C[:, None] * off + np.diag(C)
    return B, eps, beta
