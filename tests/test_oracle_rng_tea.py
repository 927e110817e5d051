"""Pins the oracle's RNG and TEA pieces: hand-derived single steps of the
published generators, the reference's own ran2.h (compiled under
oracle/_ref), Box-Muller statistics, and RPY / TEA closed forms."""
import math
import os
import subprocess

import numpy as np
import pytest

from mt_amd import params, topology
from oracle.oracle import Oracle, generate_seeds, rforce, taus_draws

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def test_taus_single_steps_by_hand():
    # One step of each component from the published recurrences
    # z' = ((z & M) << S3) ^ (((z << S1) ^ z) >> S2) and the LCG z' = a z + c:
    # state (x, y, z, w) = (0x80000000, 0x80000000, 0x80000000, 0):
    #   comp1: ((0x80000000<<13)^0x80000000)>>19 = 0x1000, (0x80000000&~1)<<12 = 0 -> 0x1000
    #   comp2: ((0x80000000<<2)^0x80000000)>>25 = 0x40,   (0x80000000&~7)<<4 = 0  -> 0x40
    #   comp3: ((0x80000000<<3)^0x80000000)>>11 = 0x100000, (..&~15)<<17 = 0      -> 0x100000
    #   LCG:   1664525*0 + 1013904223 = 0x3C6EF35F
    s = np.array([0x80000000, 0x80000000, 0x80000000, 0], np.uint32)
    out = taus_draws(s, 1)
    assert int(out[0]) == 0x1000 ^ 0x40 ^ 0x100000 ^ 0x3C6EF35F
    assert list(s) == [0x1000, 0x40, 0x100000, 0x3C6EF35F]


def test_uint_to_float_edges():
    # mantissa fill to [1,2) minus 1; an exact 0 maps to 1e-8 (src/HybridTaus.cu:53-61)
    # u = 0 in the first draw: r = sqrt(-2 log 1e-8); checked through rforce on a crafted state
    s = np.array([0x80000000, 0x80000000, 0x80000000, 0], np.uint32)
    r = rforce(s.copy())
    assert np.all(np.isfinite(r))


def test_ran2_seeds_against_reference_build():
    exe = os.path.join(REF, "ref_seeds")
    if not os.path.exists(exe):
        pytest.skip("reference ran2.h build not present (built only where the reference tree exists)")
    for seed, np_ in ((1234567, 8), (42, 100)):
        out = subprocess.run([exe, str(seed), str(np_)], capture_output=True, text=True, check=True).stdout.split()
        ref = np.array([int(v) for v in out], np.uint64).astype(np.uint32)
        np.testing.assert_array_equal(generate_seeds(seed, np_), ref)


def test_box_muller_normals_statistics():
    from scipy import stats
    s = generate_seeds(97, 1)
    draws = np.array([rforce(s) for _ in range(20000)]).ravel()
    assert abs(draws.mean()) < 0.02
    assert abs(draws.var() - 1) < 0.03
    assert stats.kstest(draws, "norm").pvalue > 1e-3


def _rpy_textbook(a, r):
    # Rotne-Prager-Yamakawa mobility relative to the self term:
    # r >= 2a: (3a/4r)[(1 + 2a^2/3r^2) I + (1 - 2a^2/r^2) e e]
    # r <  2a: (1 - 9r/32a) I + (3r/32a) e e
    if r >= 2 * a:
        return 3 * a / (4 * r) * (1 + 2 * a * a / (3 * r * r)), 3 * a / (4 * r) * (1 - 2 * a * a / (r * r))
    return 1 - 9 * r / (32 * a), 3 * r / (32 * a)


def _two_free(r, direction):
    e = np.asarray(direction, float) / np.linalg.norm(direction)
    c = np.zeros((2, 7), np.float32)
    c[1, :3] = r * e
    return topology.Batch(coords=c, off=np.array([0, 2], np.int64), harm_count=np.zeros(2, np.int32),
                          harm=np.full(2, -1, np.int32), lon_count=np.zeros(2, np.int32),
                          lon=np.zeros((2, 8), np.int32), lat_count=np.zeros(2, np.int32),
                          lat=np.zeros((2, 16), np.int32), fixed=np.zeros(2, np.uint8),
                          extra=np.zeros(2, np.uint8), mon_type=np.zeros(2, np.int32), gtp=np.ones(2, np.int32),
                          on_tub=np.zeros(2, np.int32), zs=np.zeros(1, np.float32)), e


@pytest.mark.parametrize("r", [1.5, 4.0, 9.0])
def test_rpy_epsilon_and_ci_closed_form(r):
    p = params.derive(dict(params.TEA, LJ_on=False, is_assembly=False))
    b, _ = _two_free(r, [1.0, 2.0, -0.5])
    d = b.coords[1, :3].astype(np.float64)          # the float32 geometry actually used
    r = float(np.linalg.norm(d))
    e = d / r
    o = Oracle(p, b)
    ci, eps = o.tea_eps()
    cii, crr = _rpy_textbook(p.tea_a, r)
    D = cii * np.eye(3) + crr * np.outer(e, e)
    assert eps[0] == pytest.approx(D.sum(), rel=1e-6)
    np.testing.assert_allclose(ci[0], (D ** 2).sum(1), rtol=1e-6)
    # Geyer & Winter eq. 22 and 26 for N = 2 beads
    err, eps_t, beta_t = o.tea_beta(eps)
    assert err == 0
    n = 2
    epsv = 2 * D.sum() / (3 * n * (3 * n - 3))
    aa = (3 * n - 1) * epsv ** 2 - (3 * n - 2) * epsv
    assert eps_t[0] == pytest.approx(epsv, rel=1e-9)
    assert beta_t[0] == pytest.approx((1 - math.sqrt(1 - aa)) / aa, rel=1e-6)


def test_tea_far_apart_reduces_to_free_diffusion():
    # With the beads far apart D_hat -> 0: the TEA step is plain Brownian dynamics.
    p = params.derive(dict(params.TEA, LJ_on=False, is_assembly=False, tea_epsilon_freq=1))
    b, _ = _two_free(1e5, [0, 0, 1])
    o = Oracle(p, b)
    nz = np.random.default_rng(1).standard_normal((1, 2, 6)).astype(np.float32)
    assert o.run(0, 1, noise=nz) == 0
    dx = o.b.coords[0, :3].astype(np.float64)       # bead 0 starts at the origin
    np.testing.assert_allclose(dx, p.varR * nz[0, 0, :3], rtol=2e-4, atol=1e-7)


def test_exact_em_tensor_properties():
    from oracle import em_exact
    from tests.helpers import c1_batch
    pos = c1_batch().coords[:, :3].astype(np.float64)
    D = em_exact.rpy_hat(pos, 2.0)
    assert np.abs(D - D.T).max() < 1e-12
    assert np.linalg.eigvalsh(D).min() > 0          # RPY is positive definite
    cii, crr = _rpy_textbook(2.0, float(np.linalg.norm(pos[1] - pos[0])))
    e = (pos[1] - pos[0]) / np.linalg.norm(pos[1] - pos[0])
    np.testing.assert_allclose(D[0:3, 3:6], cii * np.eye(3) + crr * np.outer(e, e), rtol=1e-12)


def test_oracle_tea_drift_is_exact_em_drift():
    # TEA's deterministic part equals the exact EM drift (dt/gammaR) D_hat F.
    from oracle import em_exact
    from tests.helpers import c1_batch
    # dt = 1000 ps makes the zero-noise step ~10 nm, so float32 positions resolve
    # the displacement to ~1e-6 (the drift is linear in dt)
    p = params.derive(dict(params.TEA, tea_epsilon_freq=1, dt=1000.0))
    b = c1_batch()
    o = Oracle(p, b.copy())
    o.rebuild()
    F = o.forces()
    c0 = b.coords.astype(np.float64)
    assert o.run(0, 1, noise=np.zeros((1, b.nb, 6), np.float32)) == 0
    mob = (b.fixed == 0) & (b.extra == 0)
    exact = em_exact.drift(c0[:, :3], F[:, :3], p.tea_a, p.dt, p.gammaR)
    d = o.b.coords[:, :3].astype(np.float64) - c0[:, :3]
    assert np.abs(d[mob] - exact[mob]).max() / np.abs(exact[mob]).max() < 5e-6
