"""Pins the MLP-HI oracle: backward against central finite differences,
the taper's closed form, and labels against a direct RPY sum."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = M.init_weights(1).astype(np.float64)
    w += rng.normal(0, 0.05, w.shape)
    x = rng.normal(0, 0.5, (7, M.IN))
    t = rng.normal(0, 0.3, (7, M.OUT))
    loss, g = M.loss_and_grad(w, x, t)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss_and_grad(wp, x, t)[0] - M.loss_and_grad(wm, x, t)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9)


def test_taper_closed_form():
    assert M.taper(5.0) == 1.0 and M.taper(12.0) == 0.0 and M.taper(15.0) == 0.0
    assert M.taper(10.0) == pytest.approx(0.5)


def test_labels_two_beads():
    # two beads 4 nm apart: A - I is the single RPY block at r = 2a
    c = np.zeros((2, 7), np.float32)
    c[1, 0] = 4.0
    feat, targ = M.local_features_targets(c, np.array([0, 2]), np.zeros(2, np.uint8), None, 2.0)
    ra = 2.0
    cii, crr = 1 - 9 * ra / 32, 3 * ra / 32
    np.testing.assert_allclose(targ[0].reshape(3, 3), np.diag([cii + crr, cii, cii]), atol=1e-12)
    np.testing.assert_allclose(feat[0, :3], [4.0 / 12.0, 0, 0])
    assert np.all(feat[0, 3:] == 0)


def test_bf16_rounding():
    # bf16 keeps 7 mantissa bits: near 1 the grid is 2^-7; halfway cases round to even
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -3.14159], np.float32)
    r = M.round_bf16(v)
    assert list(r[:4]) == [1.0, 1.0, 1 + 2 ** -6, 1.0]
    assert abs(r[4] - v[4]) <= abs(v[4]) * 2 ** -8
