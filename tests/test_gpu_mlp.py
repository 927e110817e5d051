"""GPU checks of the learned hydrodynamic correction (MLP-HI) against the
FP64 numpy oracle (oracle/mlp_oracle.py): neighbour features and exact
labels, inference with bf16-rounded weights, the training gradient, one
MLP-HI integration step with injected noise, device training and the TEA
fallback."""
import numpy as np
import pytest

from mt_amd import params as P
from mt_amd.engine import Engine, HI_MLP, MODE_STREAMED
from oracle import mlp_oracle as M
from oracle.oracle import Oracle
from tests.helpers import noise, uneven_batch

pytestmark = pytest.mark.gpu


def _system():
    p = P.derive(dict(P.TEA, tea_epsilon_freq=10))
    b = uneven_batch(seed=11, lengths=(6, 4), n_free=(10, 6))
    return p, b


def test_features_and_labels():
    print("case MLP features/labels vs oracle", flush=True)
    p, b = _system()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    f, t = eng.mlp_local()
    fo, to = M.local_features_targets(b.coords, b.off, b.extra, None, p.tea_a)
    np.testing.assert_allclose(f, fo, atol=2e-6)
    np.testing.assert_allclose(t, to, rtol=1e-5, atol=1e-5 * np.abs(to).max())
    eng.close()


def test_forward_and_gradient():
    print("case MLP forward (bf16 weights) and gradient vs oracle", flush=True)
    p, b = _system()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    w = M.init_weights(3)
    w += np.random.default_rng(1).normal(0, 0.02, w.shape).astype(np.float32)
    eng.mlp_set_weights(w)
    f, t = eng.mlp_local()
    y = eng.mlp_forward(f)
    yo = M.forward(M.round_bf16(w), f)
    assert np.abs(y - yo).max() / np.abs(yo).max() < 1e-5
    loss, g = eng.mlp_loss_grad(f[:200], t[:200])
    lo, go = M.loss_and_grad(w.astype(np.float64), f[:200], t[:200])
    assert loss == pytest.approx(lo, rel=1e-5)
    assert np.abs(g - go).max() / np.abs(go).max() < 1e-4
    eng.close()


def test_mlp_step_matches_restatement():
    print("case MLP-HI step with injected noise", flush=True)
    p, b = _system()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    w = M.init_weights(5) * 0.3
    eng.mlp_set_weights(w)
    eng.set_hi_mode(HI_MLP, check_every=1000, tolerance=1e9)
    orc = Oracle(p, b.copy())
    orc.rebuild()
    F = orc.forces()
    ci, eps = orc.tea_eps()
    err, _, beta = orc.tea_beta(eps)
    assert err == 0
    f, _ = eng.mlp_local()
    y = M.forward(M.round_bf16(w), f)
    nz = noise(np.random.default_rng(2), 1, b.nb)
    eng.set_noise(nz)
    eng.run(1)
    g = eng.get_state().astype(np.float64)
    c0 = b.coords.astype(np.float64)
    mob = (b.fixed == 0) & (b.extra == 0)
    tid = np.repeat(np.arange(b.ntraj), b.sizes())
    worst = 0.0
    for i in np.nonzero(mob)[0]:
        Y = y[i].reshape(3, 3)
        G = 0.5 * (Y + Y.T)
        bt = beta[tid[i]]
        C = 1 / np.sqrt(1 + bt * bt * ci[i])
        S = np.outer(C, C) * (np.eye(3) + 2 * bt * G)
        Lc = np.linalg.cholesky(S)
        d = p.dt / p.gammaR * (F[i, :3] + G @ F[i, :3]) + p.varR * Lc @ nz[0, i, :3]
        worst = max(worst, np.abs(g[i, :3] - c0[i, :3] - d).max())
    assert worst < 2e-5
    eng.close()


def test_training_reduces_loss_and_fallback_fires():
    print("case MLP training on device + fallback", flush=True)
    p, b = _system()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    eng.mlp_set_weights(M.init_weights(7))
    rng = np.random.default_rng(3)
    feats, labs = [], []
    for _ in range(8):
        eng.set_noise(noise(rng, 20, b.nb))
        eng.run(20)
        f, t = eng.mlp_local()
        mob = (b.fixed == 0) & (b.extra == 0)
        feats.append(f[mob])
        labs.append(t[mob])
    X, Y = np.concatenate(feats), np.concatenate(labs)
    losses = eng.mlp_train(X, Y, iters=300, batch=256, lr=2e-3)
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()
    # an untrained network must trip the 2% judge and hand the run to TEA
    eng.mlp_set_weights(M.init_weights(9))
    eng.set_noise(None)
    eng.set_hi_mode(HI_MLP, check_every=5, tolerance=0.02)
    eng.run(10)
    err, fb, bad = eng.mlp_status()
    assert fb and err.max() > 0.02
    assert np.all(np.isfinite(eng.get_state()))
    eng.close()


def test_trained_network_gains_over_free_draining():
    """What the learned correction buys: on 16 systems of a 20-dimer tube with
    2000 free dimers (periodic box at 30 uM), the free-dimer centre-of-mass
    diffusion coefficient without HI misses TEA's by several percent; the
    network, trained on the device from the engine's own TEA frames, closes
    most of that gap.  Standard error of each D is ~0.6% (32k dimers)."""
    print("case MLP gain over free draining (train + 3 x 300 steps)", flush=True)
    from mt_amd import mlp_hi, workloads
    p = P.derive(dict(P.TEA))
    b = workloads.uneven(16, dimers=(20, 20), free=(2000, 2000), conc=(30, 30), seed=23)
    X, Y, _ = mlp_hi.training_data(p, workloads.uneven(8, dimers=(20, 30), free=(1500, 2500), conc=(20, 40),
                                                       seed=29), frames=8, every=20)
    w, losses = mlp_hi.train(p, b, X, Y, iters=12000, bsize=4096, lr=2e-3)
    print(f"training loss {losses[:5].mean():.4g} -> {losses[-20:].mean():.4g}", flush=True)
    assert losses[-20:].mean() < 0.01 * losses[:5].mean()
    D = {}
    for mode in ("tea", "none", "mlp"):
        dt, s0, s1, st = mlp_hi.run_mode(p, b, mode, steps=300, warmup=10, weights=w, check_every=100,
                                         tolerance=0.05)
        D[mode], err = mlp_hi.com_diffusion(b, s0, s1, 300 * p.dt)
        assert err < 0.01 * D[mode]
        if mode == "mlp":
            errs, fb, bad = st
            print("sampled errors", np.round(errs, 4), "fallback", fb, "bad", bad, flush=True)
            assert not fb and bad == 0 and errs.max() < 0.05
    gap_none = abs(D["none"] / D["tea"] - 1)
    gap_mlp = abs(D["mlp"] / D["tea"] - 1)
    print(f"D tea {D['tea']:.5g} none {D['none']:.5g} mlp {D['mlp']:.5g}", flush=True)
    assert gap_none > 0.03
    assert gap_mlp < 0.5 * gap_none
