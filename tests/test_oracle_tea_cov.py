"""Pins the FP64 TEA noise matrix of oracle/em_exact.py (the checker of the
GPU's TEA covariance error) without running the reference.

TEA (Geyer & Winter 2009; src/bdhitea_kernel.cu:104-213, src/bdhitea.cu:57-118)
replaces the exact Cholesky factor L (L L^T = D_hat) by B with B_ii = C_i,
B_ij = beta D_hat_ij diag(C_i) (see tea_b_from).  Checked here:
  * em_exact's epsilon, beta and row sums ci equal the C oracle's TEA refresh
    (itself pinned by N = 2 closed forms in test_oracle_rng_tea.py);
  * with Geyer & Winter's row scaling every row of B has unit norm:
    C_i = 1 / sqrt(1 + beta^2 ci) makes each coordinate's variance exact
    (eq. 19); the reference scales R_j's components by C_i instead
    (src/bdhitea_kernel.cu:115-118), which moves the row norms off 1;
  * the C oracle's TEA step is the reference form of B applied to the noise:
    unit noise vectors injected one column at a time give back B;
  * TEA becomes exact as the coupling vanishes: the covariance error of B
    falls towards zero as the same tube is dilated.
"""
import numpy as np
import pytest

from mt_amd import lattice, params, topology
from oracle import em_exact
from oracle.oracle import Oracle

A = 2.0


def _tube(n_dimers=4, scale=1.0, jitter=0.05, seed=3):
    b = topology.replicate(lattice.microtubule(n_dimers), 1, jitter, seed)
    b.coords[:, :3] = (b.coords[:, :3].astype(np.float64) * scale).astype(np.float32)
    return b


def _free(b):
    """No bonded, LJ or wall forces: a TEA step moves beads by noise alone."""
    b.harm_count[:] = 0
    b.harm[:] = -1
    b.fixed[:] = 0
    return params.derive(dict(params.TEA, tea_epsilon_freq=1, LJ_on=False, is_assembly=False, barrier=False,
                              dt=1e4))


def test_em_exact_tea_refresh_equals_oracle():
    b = _tube(12)
    p = params.derive(dict(params.TEA))
    o = Oracle(p, b.copy())
    ci, eps_b = o.tea_eps()
    err, eps_t, beta_t = o.tea_beta(eps_b)
    assert err == 0
    pos = b.coords[:, :3].astype(np.float64)
    B, eps, beta = em_exact.tea_b_matrix(pos, A)
    assert eps == pytest.approx(eps_t[0], rel=1e-12)
    assert beta == pytest.approx(beta_t[0], rel=1e-7)   # the reference keeps beta in float
    off = em_exact._off_diagonal(em_exact.rpy_hat(pos, A))
    np.testing.assert_allclose((off ** 2).sum(1).reshape(-1, 3), ci, rtol=1e-12)


def test_tea_row_norms():
    pos = _tube(12).coords[:, :3].astype(np.float64)
    B, _, _ = em_exact.tea_b_matrix(pos, A, form="gw")
    np.testing.assert_allclose((B * B).sum(1), 1.0, rtol=1e-13)
    Br, _, _ = em_exact.tea_b_matrix(pos, A)
    rn = (Br * Br).sum(1)
    assert 1e-4 < np.abs(rn - 1).max() < 0.05
    # the diagonal blocks (B_ii = diag C_i) are the same in both forms
    np.testing.assert_array_equal(np.diag(Br), np.diag(B))


def test_oracle_tea_step_applies_b():
    b = _tube(2)
    p = _free(b)
    pos = b.coords[:, :3].astype(np.float64)
    B, _, _ = em_exact.tea_b_matrix(pos, A)
    n3 = 3 * b.nb
    cols = np.zeros((n3, n3))
    for k in range(n3):
        o = Oracle(p, b.copy())
        nz = np.zeros((1, b.nb, 6), np.float32)
        nz[0, k // 3, k % 3] = 1.0
        assert o.run(0, 1, noise=nz) == 0
        cols[:, k] = (o.b.coords[:, :3].astype(np.float64) - pos).reshape(-1) / p.varR
    # float32 positions of ~10 nm against displacements of varR ~ 1.6 nm
    assert np.abs(cols - B).max() < 1e-5


def test_tea_error_vanishes_with_dilution():
    errs = []
    for s in (1.0, 2.0, 4.0, 16.0):
        pos = _tube(4, scale=s).coords[:, :3].astype(np.float64)
        B, _, _ = em_exact.tea_b_matrix(pos, A)
        errs.append(em_exact.covariance_error(B, em_exact.rpy_hat(pos, A)))
    assert all(e1 > e2 for e1, e2 in zip(errs, errs[1:]))
    assert errs[0] > 0.1 and errs[-1] < 0.02
