"""The TEA pair sums on the matrix cores (tea_chunk MTE_TEA_MFMA) against the
oracle and the FP64 Ermak-McCammon tensor: TEA steps, the deterministic
drift on 312- and 1040-bead microtubules, an uneven batch with reservoir
(extra) beads, and the noise operator B recovered column by column.  The
split-bf16 products carry ~2^-17 relative error per term (amplified at most
~3x by the moment expansion beyond 16 nm); the bounds below are the FP32
ones of the VALU kernels widened to 1e-4."""
import numpy as np
import pytest

from mt_amd import lattice, params as P, topology
from mt_amd.engine import Engine, EngineError, MODE_STREAMED, TEA_MFMA
from oracle import em_exact
from oracle.oracle import Oracle
from tests.helpers import c1_batch, noise, uneven_batch

pytestmark = pytest.mark.gpu


def test_mfma_tea_steps_match_oracle():
    print("case MFMA TEA: eps/beta + 10 steps vs oracle, 312 beads", flush=True)
    p = P.derive(dict(P.TEA, tea_epsilon_freq=5))
    b = c1_batch()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    eng.set_tuning(tea_chunk=TEA_MFMA)
    assert eng.tuning()["tea_chunk"] == TEA_MFMA
    orc = Oracle(p, b.copy(), precision="f64")
    nz = noise(np.random.default_rng(5), 10, b.nb)
    eng.set_noise(nz)
    eng.run(10)
    assert orc.run(0, 10, noise=nz) == 0
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("MFMA TEA max |dx|", dx, flush=True)
    assert dx < 1e-4
    eng.close()


@pytest.mark.parametrize("n_dimers", [12, 40])
def test_mfma_tea_drift_against_fp64_exact_em(n_dimers):
    print(f"case MFMA TEA drift vs FP64 exact EM, {26 * n_dimers} beads", flush=True)
    p = P.derive(dict(P.TEA, tea_epsilon_freq=1, dt=1000.0))
    b = topology.replicate(lattice.microtubule(n_dimers), 1, 0.05, 1234567)
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    eng.set_tuning(tea_chunk=TEA_MFMA)
    orc = Oracle(p, b.copy())
    orc.rebuild()
    F = orc.forces()
    eng.set_noise(np.zeros((1, b.nb, 6), np.float32))
    c0 = b.coords.astype(np.float64)
    eng.run(1)
    g = eng.get_state().astype(np.float64)
    eng.close()
    mob = (b.fixed == 0) & (b.extra == 0)
    exact = em_exact.drift(c0[:, :3], F[:, :3], p.tea_a, p.dt, p.gammaR)
    d = g[:, :3] - c0[:, :3]
    err = np.abs(d[mob] - exact[mob]).max() / np.abs(exact[mob]).max()
    print("MFMA TEA drift relative error vs FP64", err, flush=True)
    assert err < 1e-4


def test_mfma_tea_uneven_batch_with_extra_beads():
    print("case MFMA TEA uneven batch with extra beads", flush=True)
    p = P.derive(dict(P.TEA, tea_epsilon_freq=3))
    b = uneven_batch()
    s, e = int(b.off[1]), int(b.off[2])
    b.extra[s + 13 * 2 * 5:e] = 1
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    eng.set_tuning(tea_chunk=TEA_MFMA)
    orc = Oracle(p, b.copy(), precision="f64")
    nz = noise(np.random.default_rng(9), 6, b.nb)
    eng.set_noise(nz)
    eng.run(6)
    assert orc.run(0, 6, noise=nz) == 0
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("MFMA TEA uneven max |dx|", dx, flush=True)
    assert dx < 1e-4
    eng.close()


def test_mfma_tea_noise_operator():
    print("case MFMA TEA noise operator B vs FP64, 312 beads", flush=True)
    ncopy = 64
    one = topology.replicate(lattice.microtubule(12), 1, 0.05, 1234567)
    b = topology.replicate(lattice.microtubule(12), ncopy, 0.0, 1)
    b.coords = np.tile(one.coords, (ncopy, 1))
    b.harm_count[:] = 0
    b.harm[:] = -1
    b.fixed[:] = 0
    p = P.derive(dict(P.TEA, tea_epsilon_freq=1, LJ_on=False, is_assembly=False, barrier=False, dt=1e4))
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_STREAMED)
    eng.set_tuning(tea_chunk=TEA_MFMA)
    n = one.nb
    n3 = 3 * n
    Bg = np.zeros((n3, n3))
    for c0 in range(0, n3, ncopy):
        nz = np.zeros((1, b.nb, 6), np.float32)
        ks = list(range(c0, min(c0 + ncopy, n3)))
        for t, k in enumerate(ks):
            nz[0, t * n + k // 3, k % 3] = 1.0
        eng.set_state(b.coords)
        eng.step = 0
        eng.set_noise(nz)
        eng.run(1)
        d = (eng.get_state()[:, :3].astype(np.float64) - b.coords[:, :3].astype(np.float64)).reshape(ncopy, -1)
        for t, k in enumerate(ks):
            Bg[:, k] = d[t] / p.varR
    _, beta = eng.tea_stats()
    ci = eng.tea_ci()[:n]
    eng.close()
    pos = one.coords[:, :3].astype(np.float64)
    Dh = em_exact.rpy_hat(pos, 2.0)
    B = em_exact.tea_b_from(Dh, beta[0], ci)
    rel = np.linalg.norm(Bg - B) / np.linalg.norm(B)
    print(f"MFMA B vs FP64 B {rel:.2e}; covariance error {em_exact.covariance_error(Bg, Dh):.5f}", flush=True)
    assert rel < 5e-5


def test_mfma_tea_rejected_for_periodic_boxes():
    p = P.derive(dict(P.TEA))
    b = uneven_batch()
    b.box = np.array([130.0, 100.0, 180.0], np.float32)
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    with pytest.raises(EngineError) as ei:
        eng.set_tuning(tea_chunk=TEA_MFMA)
    assert ei.value.code == 1
    eng.close()


def test_mfma_tea_rejected_above_spatial_order_limit():
    # 81 dimers x 26 beads = 2106 > 2048: k_rebuild_order keeps the index order
    p = P.derive(dict(P.TEA))
    b = topology.replicate(lattice.microtubule(81), 1, 0.05, 7)
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    with pytest.raises(EngineError) as ei:
        eng.set_tuning(tea_chunk=TEA_MFMA)
    assert ei.value.code == 1
    eng.close()
