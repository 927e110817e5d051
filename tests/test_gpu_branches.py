"""GPU parity for the force-field branches the default workload leaves idle,
TEA on the engine's own HybridTaus streams, and the error paths.

Branches (against the FP64 oracle, 1e-5 of the largest component of the same
kind, each term family alone and all together):
  * bending offsets theta0_gdp / theta0_gtp / psi_0 / fi_0 all non-zero, with
    GTP and GDP beads mixed (src/compute_cuda.cu:156-177, 271-295, 751-757,
    813-817), in both R_MON branches;
  * the barrier only between off-tubule beads, flags mixed (:259, :454);
  * seam weakening with seam_coeff = 2 (:447-452, :875-880);
  * cylindrical and end walls (:497-517);
  * the REGULARIZATION cap on LJ pair forces (:482-486) as a run-time switch.
Trajectories of 300 steps with injected noise run against the FP32 oracle in
streamed and resident modes.
"""
import numpy as np
import pytest

from mt_amd.engine import Engine, EngineError, MODE_RESIDENT, MODE_STREAMED
from oracle.oracle import Oracle, generate_seeds
from tests.helpers import ANG, XYZ, branch_batch, branch_params, noise, rel_err

pytestmark = pytest.mark.gpu

# Parameter overrides that leave one term family alone.  Harmonic partners
# are removed (harm_count 0) where the harmonic term must vanish.
TERMS = {
    "all": (dict(), True),
    "harmonic": (dict(is_assembly=False, LJ_on=False, repulsive_walls=False), True),
    "longitudinal": (dict(D_lat=0.0, a_barr_lat=0.0, LJ_on=False, repulsive_walls=False), False),
    "lateral": (dict(D_long=0.0, a_barr_long=0.0, B_psi=0.0, B_fi=0.0, B_theta=0.0, LJ_on=False,
                     repulsive_walls=False), False),
    "lj": (dict(is_assembly=False, repulsive_walls=False), False),
    "walls": (dict(is_assembly=False, LJ_on=False, rep_eps=5.0), False),
}


def _check_lists(eng, orc):
    L = eng.lists()
    b = orc.b
    np.testing.assert_array_equal(L["lon_count"], b.lon_count)
    np.testing.assert_array_equal(L["lat_count"], b.lat_count)
    np.testing.assert_array_equal(L["lj_count"], orc.lj_count)
    for k in range(8):
        m = b.lon_count > k
        np.testing.assert_array_equal(L["lon"][m, k], b.lon[m, k])
    for k in range(16):
        m = b.lat_count > k
        np.testing.assert_array_equal(L["lat"][m, k], b.lat[m, k])


@pytest.mark.parametrize("term", list(TERMS))
def test_branch_forces_energies(term):
    over, harmonic = TERMS[term]
    print(f"case branch forces/energies: {term}", flush=True)
    p = branch_params(**over)
    b = branch_batch(ntraj=2, jitter=0.2)
    if not harmonic:
        b.harm_count[:] = 0
        b.harm[:] = -1
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_STREAMED)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f64")
    eng.rebuild()
    assert orc.rebuild() == 0
    _check_lists(eng, orc)
    f = eng.forces()
    fo = orc.forces()
    assert np.abs(fo[:, XYZ]).max() > 0
    assert rel_err(f, fo, XYZ) < 1e-5, term
    if np.abs(fo[:, ANG]).max() > 0:
        assert rel_err(f, fo, ANG) < 1e-5, term
    eb, et = eng.energies()
    eo = orc.energies()
    for q in range(7):
        s = np.abs(eo[:, q]).max()
        if s > 0:
            assert np.abs(eb[:, q] - eo[:, q]).max() / s < 1e-5, (term, q)
    if term == "all":
        # the branches are live in this input: barrier terms, seam bonds,
        # GDP and GTP offsets and every wall
        off = Oracle(branch_params(barrier=False), orc.b.copy(), max_lj=64)
        off.lj_count[:], off.lj[:] = orc.lj_count, orc.lj
        assert np.abs(off.forces() - fo).max() > 1e-3 * np.abs(fo).max()
        assert (orc.b.mon_type != orc.b.mon_type[0]).any() and set(np.unique(orc.b.gtp)) == {0, 1}
        x = orc.b.coords
        assert (x[:, 2] < p.rep_leftborder).any() and (x[:, 2] > 40.0 + p.rep_leftborder).any()
        assert (np.hypot(x[:, 0], x[:, 1]) > p.rep_r).any()
    eng.close()


@pytest.mark.parametrize("mode", [MODE_STREAMED, MODE_RESIDENT])
def test_branch_trajectory_injected_noise(mode):
    print(f"case branch trajectory 300 steps, mode {mode}", flush=True)
    p = branch_params()
    b = branch_batch(ntraj=2, jitter=0.2)
    eng = Engine(p, b.copy(), max_lj=64, mode=mode)
    assert eng.mode == mode
    orc = Oracle(p, b.copy(), max_lj=64, precision="f32")
    nz = noise(np.random.default_rng(5), 300, b.nb)
    eng.set_noise(nz)
    eng.run(300)
    assert orc.run(0, 300, noise=nz) == 0
    g = eng.get_state()
    o = orc.b.coords
    dx = np.abs(g[:, :3] - o[:, :3]).max()
    da = np.abs(g[:, 3:6] - o[:, 3:6]).max()
    print("max |dx| nm", dx, "max |dang| rad", da, flush=True)
    assert dx < 1e-3 and da < 1e-3
    eng.close()


def _overlap_batch():
    """uneven_batch with trajectory 1's first reservoir dimer moved rigidly so
    that its bead k sits 0.9 nm radially outside the mobile lattice bead i
    (its partner lands ~0.9 nm from i's upper neighbour): LJ pairs well
    inside the range where the REGULARIZATION cap binds (~1.02 nm at
    dt = 200 ps)."""
    from tests.helpers import uneven_batch
    b = uneven_batch()
    s = int(b.off[1])
    i, k = s + 128, s + 130
    assert b.extra[k] and b.extra[k + 1] and not b.extra[i] and not b.fixed[i]
    x = b.coords[i, :3].astype(np.float64)
    u = np.array([x[0], x[1], 0.0]) / np.hypot(x[0], x[1])
    shift = x + 0.9 * u - b.coords[k, :3].astype(np.float64)
    b.coords[k:k + 2, :3] = (b.coords[k:k + 2, :3].astype(np.float64) + shift).astype(np.float32)
    return b, i


@pytest.mark.parametrize("mode", [MODE_STREAMED, MODE_RESIDENT])
def test_lj_regularization(mode):
    """REGULARIZATION (src/compute_cuda.cu:482-486) as the run-time switch:
    forces against the FP64 oracle with the cap binding on two pairs, and 50
    steps of injected noise against the FP32 oracle, in both modes."""
    print(f"case LJ regularization, mode {mode}", flush=True)
    from tests.helpers import base_params
    b, i = _overlap_batch()
    p = base_params(regularization=True)
    cap = 2.0 * float(p.gammaR) / float(p.dt)
    off = Oracle(base_params(), b.copy(), max_lj=64, precision="f64")
    orc = Oracle(p, b.copy(), max_lj=64, precision="f64")
    assert off.rebuild() == 0 and orc.rebuild() == 0
    fo, f_off = orc.forces(), off.forces()
    # the cap binds: bead i's unclamped push is ~2.4x the cap
    assert np.linalg.norm(f_off[i, :3] - fo[i, :3]) > cap
    eng = Engine(p, b.copy(), max_lj=64, mode=mode)
    assert eng.mode == mode
    if mode == MODE_STREAMED:
        eng.rebuild()
        assert rel_err(eng.forces(), fo, XYZ) < 1e-5
    o32 = Oracle(p, b.copy(), max_lj=64, precision="f32")
    nz = noise(np.random.default_rng(8), 50, b.nb)
    eng.set_noise(nz)
    eng.run(50)
    assert o32.run(0, 50, noise=nz) == 0
    dx = np.abs(eng.get_state()[:, :3] - o32.b.coords[:, :3]).max()
    print("max |dx| nm", dx, flush=True)
    assert dx < 1e-3
    eng.close()


@pytest.mark.parametrize("mode", [MODE_STREAMED, MODE_RESIDENT])
def test_branch_rng_streams(mode):
    """Mobile beads draw both streams every step, fixed and extra beads none
    (src/compute_cuda.cu:949-953), in both modes."""
    print(f"case branch HybridTaus streams, mode {mode}", flush=True)
    from tests.helpers import uneven_batch
    p = branch_params()
    b = uneven_batch()
    eng = Engine(p, b.copy(), max_lj=64, mode=mode)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f32")
    seeds = generate_seeds(1234567, 2 * b.nb)
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    eng.run(30)
    assert orc.run(0, 30, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    assert np.abs(eng.get_state()[:, :6] - orc.b.coords[:, :6]).max() < 1e-3
    eng.close()


def test_tea_hybridtaus_draw_order():
    """TEA draws both HybridTaus streams for every bead, fixed and extra ones
    included (integrateTea_prepare and integrateTea_kernel_unlisted,
    src/bdhitea_kernel.cu:22,194): after 12 TEA steps on the engine's own
    streams every state word equals the oracle's, and the positions agree
    to the FP32 Box-Muller rounding."""
    print("case TEA on HybridTaus streams, uneven batch with extra beads", flush=True)
    from mt_amd import params as P
    from tests.helpers import uneven_batch
    p = P.derive(dict(P.TEA, tea_epsilon_freq=5))
    b = uneven_batch()
    assert b.extra.any() and b.fixed.any()
    eng = Engine(p, b.copy(), max_lj=64)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f64")
    seeds = generate_seeds(1234567, 2 * b.nb)
    eng.run(12)
    assert orc.run(0, 12, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    # the plain integrator would have left the extra beads' streams alone
    untouched = generate_seeds(1234567, 2 * b.nb).reshape(-1, 4)
    ex = np.flatnonzero(b.extra)
    ex = np.concatenate([ex, ex + b.nb])   # both streams
    assert (seeds.reshape(-1, 4)[ex] != untouched[ex]).all()
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("TEA max |dx|", dx, flush=True)
    assert dx < 1e-4
    eng.close()


# ------------------------------------------------------------------ errors
@pytest.mark.parametrize("mode", [MODE_STREAMED, MODE_RESIDENT])
def test_list_overflow_stops_before_stepping(mode):
    """An LJ list capacity below the ~12 neighbours within 8 nm overflows at
    the first rebuild.  Every kernel exits normally; mte_run reports
    MTE_ERR_OVERFLOW and leaves the state, the RNG streams and the step
    counter at that rebuild's step (src/compute_cuda.cu:647-660 writes past
    the capacity instead)."""
    print(f"case list overflow, max_lj 4, mode {mode}", flush=True)
    from tests.helpers import base_params, c1_batch
    p = base_params()
    b = c1_batch(ntraj=2)
    eng = Engine(p, b.copy(), max_lj=4, mode=mode)
    x0, s0 = eng.get_state(), eng.get_rng()
    with pytest.raises(EngineError) as ei:
        eng.run(250)
    assert ei.value.code == 3 and "step 0" in str(ei.value)
    assert eng.step == 0
    np.testing.assert_array_equal(eng.get_state(), x0)
    np.testing.assert_array_equal(eng.get_rng(), s0)
    # the next call rebuilds again and fails the same way
    with pytest.raises(EngineError) as ei:
        eng.run(10)
    assert ei.value.code == 3
    eng.close()
    # an overflow at a later rebuild: 100 good steps, then the lists of step
    # 100 overflow after the beads are packed closer
    eng = Engine(p, b.copy(), max_lj=24, mode=mode)
    eng.run(100)
    x1 = eng.get_state()
    packed = x1.copy()
    packed[:, :3] *= np.float32(0.5)
    eng.set_state(packed)
    with pytest.raises(EngineError) as ei:
        eng.run(150)
    assert ei.value.code == 3 and "step 100" in str(ei.value)
    assert eng.step == 100
    np.testing.assert_array_equal(eng.get_state(), packed)
    eng.close()


def test_tea_capricious_errors():
    """src/bdhitea.cu:89-107: epsilon above tea_epsmax stops the run; so does
    |a| < 1e-7 with tea_capricious on (beads far apart), while with it off
    beta falls back to 0.5 and the run goes on."""
    print("case TEA capricious checks", flush=True)
    from mt_amd import params as P
    from tests.helpers import c1_batch
    b = c1_batch()
    eng = Engine(P.derive(dict(P.TEA, tea_epsmax=1e-6)), b.copy(), max_lj=64)
    with pytest.raises(EngineError) as ei:
        eng.run(5)
    assert ei.value.code == 4
    eng.close()
    # free beads 1e9 nm apart on a line (no bonds): epsilon ~ 1e-11, |a| < 1e-7
    far = b.copy()
    far.harm_count[:] = 0
    far.harm[:] = -1
    far.coords[:, :3] = 0.0
    far.coords[:, 0] = np.arange(far.nb, dtype=np.float32) * np.float32(1e9)
    for capricious, code in ((True, 4), (False, None)):
        eng = Engine(P.derive(dict(P.TEA, tea_capricious=capricious, LJ_on=False, is_assembly=False)), far.copy(),
                     max_lj=64)
        if code:
            with pytest.raises(EngineError) as ei:
                eng.run(5)
            assert ei.value.code == code
        else:
            eng.run(5)
            _, beta = eng.tea_stats()
            assert np.all(beta == 0.5)
        eng.close()
