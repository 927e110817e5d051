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
B_ii = C_i, B_ij = beta D_hat_ij C_i (src/bdhitea_kernel.cu:104-213; see
tea_b_from for where C_i acts).
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
    """One exact Cholesky EM step; xi: (N, 3) standard normals, or (K, N, 3)
    for K independent steps from the same state."""
    Dh = rpy_hat(pos, a)
    L = np.linalg.cholesky(Dh)
    det = dt / gammaR * (Dh @ np.asarray(F, np.float64).reshape(-1))
    xi = np.asarray(xi, np.float64)
    sto = np.sqrt(2 * KB * temp * dt / gammaR) * (xi.reshape(-1, L.shape[0]) @ L.T)
    return (np.asarray(pos, np.float64).reshape(-1) + det + sto).reshape(xi.shape)


def _off_diagonal(Dh):
    off = Dh.copy()
    for i in range(0, Dh.shape[0], 3):
        off[i:i + 3, i:i + 3] = 0
    return off


def tea_b_matrix(pos, a, form="reference"):
    """TEA's noise matrix B (3N, 3N) with its FP64 epsilon and beta (Geyer &
    Winter 2009, eqs. 22, 26): epsilon the mean off-diagonal entry of D_hat,
    beta from a = (3N-1) eps^2 - (3N-2) eps, C_i per coordinate from the row
    sums of squares ci (eq. 19).  See tea_b_from for the two forms."""
    Dh = rpy_hat(pos, a)
    n3 = Dh.shape[0]
    off = _off_diagonal(Dh)
    eps = off.sum() / (n3 * (n3 - 3))
    aa = (n3 - 1) * eps * eps - (n3 - 2) * eps
    beta = 0.5 if abs(aa) < 1e-7 else (1 - np.sqrt(1 - aa)) / aa
    ci = (off ** 2).sum(1)
    return tea_b_from(Dh, beta, ci, form), eps, beta


def tea_b_from(Dh, beta, ci, form="reference"):
    """B in FP64 from D_hat and a given beta and per-row ci (3N values, the
    x, y, z rows of each bead), e.g. the values a GPU refresh produced, with
    C = 1 / sqrt(1 + beta^2 ci) and B_ii = diag(C_i).

    form "reference": B_ij = beta D_hat_ij diag(C_i).  The reference scales
    the components of R_j by bead i's C_i before the tensor product
    (integrateTea_force, src/bdhitea_kernel.cu:115-118 with
    ci = beta C_i from :171-176), so C_i acts on the columns of D_hat_ij.
    form "gw": B_ij = beta diag(C_i) D_hat_ij, Geyer & Winter's row scaling,
    which gives every row of B unit norm.  The two agree when the three
    components of C_i are equal."""
    off = _off_diagonal(Dh)
    C = 1.0 / np.sqrt(1 + beta * beta * np.asarray(ci, np.float64).reshape(-1))
    if form == "gw":
        return beta * C[:, None] * off + np.diag(C)
    if form != "reference":
        raise ValueError(form)
    n3 = len(C)
    rows = np.arange(n3) // 3 * 3                   # first coordinate of row bead i
    fac = C[rows[:, None] + (np.arange(n3) % 3)[None, :]]   # C_i of the column's component
    return beta * off * fac + np.diag(C)


def covariance_error(B, Dh):
    """Relative Frobenius error of TEA's noise covariance B B^T against the
    exact one, D_hat = L L^T (the normalised diffusion tensor)."""
    return float(np.linalg.norm(B @ B.T - Dh) / np.linalg.norm(Dh))
