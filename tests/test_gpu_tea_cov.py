"""TEA's noise covariance on the GPU against the FP64 Cholesky
Ermak-McCammon oracle, on the 312-bead microtubule (13 PF x 12 dimers).

With no forces a TEA step moves the beads by varR B xi alone
(src/bdhitea_kernel.cu:104-213), so:
  * injected unit noise vectors recover the GPU's B column by column; it
    equals the FP64 B built from the GPU's own beta and ci, and the TEA
    covariance error |B B^T - D_hat|_F / |D_hat|_F (D_hat = L L^T, the exact
    covariance) follows from it;
  * displacements drawn from the engine's own HybridTaus streams have the
    covariance B B^T within sampling error, measured against em_step draws
    of the exact (Cholesky) scheme on the same entries.
64 copies of the tube run side by side, one column or draw each per call.
"""
import numpy as np
import pytest

from mt_amd import lattice, params, topology
from mt_amd.engine import Engine, MODE_STREAMED
from oracle import em_exact

pytestmark = pytest.mark.gpu

NCOPY = 64
A = 2.0


def _setup():
    one = topology.replicate(lattice.microtubule(12), 1, 0.05, 1234567)
    b = topology.replicate(lattice.microtubule(12), NCOPY, 0.0, 1)
    b.coords = np.tile(one.coords, (NCOPY, 1))
    b.harm_count[:] = 0
    b.harm[:] = -1
    b.fixed[:] = 0
    # no forces; dt 1e4 ps makes varR ~ 1.6 nm, resolved to ~5e-6 by float32 positions
    p = params.derive(dict(params.TEA, tea_epsilon_freq=1, LJ_on=False, is_assembly=False, barrier=False, dt=1e4))
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_STREAMED)
    return p, b, one, eng


def _step(eng, b, noise=None):
    eng.set_state(b.coords)
    eng.step = 0
    if noise is not None:
        eng.set_noise(noise)
    eng.run(1)
    d = eng.get_state()[:, :3].astype(np.float64) - b.coords[:, :3].astype(np.float64)
    return d.reshape(NCOPY, -1)


def test_tea_noise_operator_and_covariance_error():
    print("case TEA noise operator B on the GPU vs FP64, 312 beads", flush=True)
    p, b, one, eng = _setup()
    n = one.nb
    n3 = 3 * n
    Bg = np.zeros((n3, n3))
    for c0 in range(0, n3, NCOPY):
        nz = np.zeros((1, b.nb, 6), np.float32)
        ks = list(range(c0, min(c0 + NCOPY, n3)))
        for t, k in enumerate(ks):
            nz[0, t * n + k // 3, k % 3] = 1.0
        d = _step(eng, b, nz)
        for t, k in enumerate(ks):
            Bg[:, k] = d[t] / p.varR
    eps, beta = eng.tea_stats()
    ci = eng.tea_ci()[:n]
    eng.close()
    pos = one.coords[:, :3].astype(np.float64)
    Dh = em_exact.rpy_hat(pos, A)
    B_own = em_exact.tea_b_from(Dh, beta[0], ci)
    B_exact_refresh, eps64, beta64 = em_exact.tea_b_matrix(pos, A)
    assert eps[0] == pytest.approx(eps64, rel=1e-5) and beta[0] == pytest.approx(beta64, rel=1e-5)
    rel = np.linalg.norm(Bg - B_own) / np.linalg.norm(B_own)
    rel64 = np.linalg.norm(Bg - B_exact_refresh) / np.linalg.norm(B_exact_refresh)
    err_gpu = em_exact.covariance_error(Bg, Dh)
    err_64 = em_exact.covariance_error(B_exact_refresh, Dh)
    var_err = float(np.abs((Bg * Bg).sum(1) - 1.0).max())
    print(f"B_gpu vs FP64 B (GPU beta, ci) {rel:.2e}, vs all-FP64 B {rel64:.2e}; covariance error "
          f"GPU {err_gpu:.5f} FP64 {err_64:.5f}; per-coordinate variance error {var_err:.2e}", flush=True)
    assert rel < 2e-5 and rel64 < 5e-5
    assert abs(err_gpu - err_64) < 1e-4
    assert var_err < 1e-3


def test_tea_sampled_covariance_from_hybridtaus():
    print("case TEA displacement covariance from HybridTaus draws, 4096 samples", flush=True)
    p, b, one, eng = _setup()
    n = one.nb
    X = np.concatenate([_step(eng, b) for _ in range(64)]) / p.varR      # (4096, 3n)
    eps, beta = eng.tea_stats()
    ci = eng.tea_ci()[:n]
    eng.close()
    pos = one.coords[:, :3].astype(np.float64)
    Dh = em_exact.rpy_hat(pos, A)
    B = em_exact.tea_b_from(Dh, beta[0], ci)
    S_tea = B @ B.T
    K = len(X)
    S = X.T @ X / K
    rng = np.random.default_rng(17)
    E = (em_exact.em_step(pos, np.zeros_like(pos), rng.standard_normal((K, n, 3)), A, p.dt, p.gammaR, p.temp)
         - pos).reshape(K, -1) / p.varR
    S_ex = E.T @ E / K
    # entries of beads within 8 nm (self blocks included)
    r = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    near = np.kron(r < 8.0, np.ones((3, 3), bool))

    def z2(Ssamp, Sigma):
        var = (np.outer(np.diag(Sigma), np.diag(Sigma)) + Sigma ** 2) / K
        return float(np.mean((Ssamp - Sigma)[near] ** 2 / var[near]))

    z_tea, z_exact_vs_tea = z2(S, S_tea), z2(S_ex, S_tea)
    z_exact, z_tea_vs_exact = z2(S_ex, Dh), z2(S, Dh)
    print(f"mean z^2 on {int(near.sum())} near entries: GPU vs TEA {z_tea:.3f}, exact draws vs exact {z_exact:.3f}; "
          f"GPU vs exact {z_tea_vs_exact:.2f}, exact draws vs TEA {z_exact_vs_tea:.2f}", flush=True)
    assert z_tea < 1.3 and z_exact < 1.3
    # the TEA approximation error is resolved by 4096 draws on these entries
    assert z_tea_vs_exact > 10.0 and z_exact_vs_tea > 10.0
    assert np.abs(np.diag(S) - 1).max() < 6 * np.sqrt(2.0 / K)
