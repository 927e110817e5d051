"""GPU parity: the HIP engine against the CPU oracle on identical inputs.

Tolerances (FP32 engine vs FP64 oracle): forces and energies 1e-5 relative
to the largest component of the same kind; lists bit-exact; trajectories with
injected identical noise compared against the FP32 oracle.
"""
import numpy as np
import pytest

from mt_amd.engine import Engine, MODE_STREAMED, TEA_EPS_ONEPASS, TEA_EPS_TWOSIDED, TEA_SYMMETRIC
from oracle.oracle import Oracle
from tests.helpers import ANG, XYZ, base_params, c1_batch, noise, rel_err, uneven_batch

pytestmark = pytest.mark.gpu


def _pair(p, b, mode=MODE_STREAMED, precision="f64", max_lj=64):
    eng = Engine(p, b.copy(), max_lj=max_lj, mode=mode)
    orc = Oracle(p, b.copy(), max_lj=max_lj, precision=precision)
    return eng, orc


def _lists_equal(eng, orc):
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
        # a static (PDB) lateral entry 0 is bead 0 in the j <= 0 orientation
        # (src/compute_cuda.cu:306-308), which the engine reports as -ZERO
        want = np.where(b.lat[m, k] == 0, -999999, b.lat[m, k])
        np.testing.assert_array_equal(L["lat"][m, k], want)
    for k in range(orc.max_lj):
        m = orc.lj_count > k
        np.testing.assert_array_equal(L["lj"][m, k], orc.lj[m, k])


@pytest.mark.parametrize("maker", [c1_batch, uneven_batch])
def test_lists_forces_energies(maker):
    print(f"case lists/forces/energies {maker.__name__}", flush=True)
    p = base_params()
    b = maker()
    eng, orc = _pair(p, b)
    eng.rebuild()
    assert orc.rebuild() == 0
    _lists_equal(eng, orc)
    f = eng.forces()
    fo = orc.forces()
    assert rel_err(f, fo, XYZ) < 1e-5
    assert rel_err(f, fo, ANG) < 1e-5
    eb, et = eng.energies()
    eo = orc.energies()
    for q in range(7):
        s = np.abs(eo[:, q]).max()
        if s > 0:
            assert np.abs(eb[:, q] - eo[:, q]).max() / s < 1e-5, q
    eng.close()


def test_trajectory_injected_noise():
    print("case trajectory 1000 steps injected noise", flush=True)
    p = base_params()
    b = c1_batch()
    eng, orc = _pair(p, b, precision="f32")
    rng = np.random.default_rng(11)
    nsteps = 1000
    nz = noise(rng, nsteps, b.nb)
    eng.set_noise(nz)
    eng.run(nsteps)
    assert orc.run(0, nsteps, noise=nz) == 0
    g = eng.get_state()
    o = orc.b.coords
    dx = np.abs(g[:, :3] - o[:, :3]).max()
    da = np.abs(g[:, 3:6] - o[:, 3:6]).max()
    print("max |dx| nm", dx, "max |dang| rad", da, flush=True)
    assert dx < 1e-3 and da < 1e-3
    eng.close()


def test_rng_streams_bit_exact():
    print("case HybridTaus streams", flush=True)
    from oracle.oracle import generate_seeds
    p = base_params()
    b = c1_batch()
    eng, orc = _pair(p, b, precision="f32")
    s0 = eng.get_rng()
    np.testing.assert_array_equal(s0, generate_seeds(1234567, 2 * b.nb))
    seeds = s0.copy()
    eng.run(20)
    assert orc.run(0, 20, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    g = eng.get_state()
    assert np.abs(g[:, :3] - orc.b.coords[:, :3]).max() < 1e-3
    eng.close()


@pytest.mark.parametrize("chunk", [0, TEA_SYMMETRIC])
def test_tea_step_parity(chunk):
    print(f"case TEA eps/beta + steps, tea_chunk={chunk}", flush=True)
    from mt_amd import params as P
    p = P.derive(dict(P.TEA, tea_epsilon_freq=5))
    b = c1_batch()
    eng, orc = _pair(p, b, precision="f64")
    if chunk:
        eng.set_tuning(tea_chunk=chunk)
    rng = np.random.default_rng(5)
    nsteps = 10
    nz = noise(rng, nsteps, b.nb)
    eng.set_noise(nz)
    eng.run(nsteps)
    assert orc.run(0, nsteps, noise=nz) == 0
    eps, beta = eng.tea_stats()
    assert abs(eps[0] - orc.eps_t[0]) / orc.eps_t[0] < 1e-5
    assert abs(beta[0] - orc.beta_t[0]) / orc.beta_t[0] < 1e-5
    g = eng.get_state()
    dx = np.abs(g[:, :3] - orc.b.coords[:, :3]).max()
    print("TEA max |dx|", dx, flush=True)
    assert dx < 1e-4
    eng.close()


def test_resident_injected_noise_matches_oracle():
    print("case resident 300 steps injected noise, 2 trajectories", flush=True)
    from mt_amd.engine import MODE_RESIDENT
    p = base_params()
    b = c1_batch(ntraj=2)
    eng, orc = _pair(p, b, mode=MODE_RESIDENT, precision="f32")
    assert eng.mode == MODE_RESIDENT
    rng = np.random.default_rng(21)
    nsteps = 300
    nz = noise(rng, nsteps, b.nb)
    eng.set_noise(nz)
    eng.run(nsteps)
    assert orc.run(0, nsteps, noise=nz) == 0
    g = eng.get_state()
    o = orc.b.coords
    dx = np.abs(g[:, :3] - o[:, :3]).max()
    da = np.abs(g[:, 3:6] - o[:, 3:6]).max()
    print("resident max |dx|", dx, "max |dang|", da, flush=True)
    assert dx < 1e-3 and da < 1e-3
    _lists_equal(eng, orc)
    eng.close()


def test_resident_rng_and_uneven_batch():
    print("case resident HybridTaus, uneven batch, 250 steps", flush=True)
    from mt_amd.engine import MODE_RESIDENT
    p = base_params()
    b = uneven_batch()
    eng, orc = _pair(p, b, mode=MODE_RESIDENT, precision="f32")
    assert eng.mode == MODE_RESIDENT
    seeds = eng.get_rng().copy()
    eng.run(250)
    assert orc.run(0, 250, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    g = eng.get_state()
    dx = np.abs(g[:, :3] - orc.b.coords[:, :3]).max()
    print("resident uneven max |dx|", dx, flush=True)
    assert dx < 1e-3
    eng.close()


def test_resident_two_beads_per_thread_rng_matches_oracle():
    """1040-bead trajectories with every bead mobile run two beads per thread
    on 12 waves; the second-pass beads of the first SIMD's waves (0 and 4)
    draw their noise through helper threads one step ahead: the HybridTaus
    states after the run and the positions equal the oracle's from the same
    seeds."""
    print("case resident 2 x 1040 unfixed beads, two per thread, drawn noise, 150 steps", flush=True)
    from mt_amd.engine import MODE_RESIDENT
    p = base_params()
    b = c1_batch(ntraj=2, n_dimers=40)
    b.fixed[:] = 0   # the 17th pass then holds mobile beads: the helpers draw
    eng, orc = _pair(p, b, mode=MODE_RESIDENT, precision="f32")
    assert eng.mode == MODE_RESIDENT and eng.tuning()["resident_per_thread"] == 2
    seeds = eng.get_rng().copy()
    eng.run(150)
    assert orc.run(0, 150, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("resident 2 per thread max |dx|", dx, flush=True)
    assert dx < 1e-3
    eng.close()


def test_resident_one_bead_per_thread_immobile_tail_matches_oracle():
    """1040-bead trajectories whose 26 fixed beads take the lanes past the
    1024-thread workgroup run one bead per thread by default; the fixed tail
    is only staged in LDS.  The HybridTaus states and positions equal the
    oracle's, and the two-beads-per-thread plan gives the same states bit for
    bit (each bead's sums do not depend on its lane or plan)."""
    print("case resident 2 x 1040 beads, one per thread with a fixed tail, 150 steps", flush=True)
    from mt_amd.engine import MODE_RESIDENT
    p = base_params()
    b = c1_batch(ntraj=2, n_dimers=40)
    eng, orc = _pair(p, b, mode=MODE_RESIDENT, precision="f32")
    assert eng.mode == MODE_RESIDENT and eng.tuning()["resident_per_thread"] == 1
    two = Engine(p, b.copy(), max_lj=64, mode=MODE_RESIDENT)
    two.set_tuning(resident_per_thread=2)
    seeds = eng.get_rng().copy()
    np.testing.assert_array_equal(two.get_rng(), seeds)
    eng.run(150)
    two.run(150)
    assert orc.run(0, 150, seeds=seeds) == 0
    np.testing.assert_array_equal(eng.get_rng(), seeds)
    np.testing.assert_array_equal(two.get_rng(), seeds)
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    d12 = np.abs(eng.get_state() - two.get_state()).max()
    print("resident 1 per thread max |dx|", dx, "vs two per thread", d12, flush=True)
    assert dx < 1e-3
    np.testing.assert_array_equal(eng.get_state(), two.get_state())
    eng.close()
    two.close()


def test_resident_plan_follows_the_mobile_count():
    """Unfixed 1040-bead trajectories with 30 extra beads each fit one bead
    per thread; clearing the extra flags leaves 1040 mobile beads, and the
    engine moves to two beads per thread: the run then equals that of an
    engine created with the cleared flags, bit for bit, which follows the
    oracle."""
    print("case resident plan after set_flags, 2 x 1040 unfixed beads, 60 steps", flush=True)
    from mt_amd.engine import MODE_RESIDENT
    p = base_params()
    b = c1_batch(ntraj=2, n_dimers=40)
    b.fixed[:] = 0
    be = b.copy()
    for t in range(2):
        be.extra[int(b.off[t]) + 1000:int(b.off[t]) + 1030] = 1
    eng = Engine(p, be, max_lj=64, mode=MODE_RESIDENT)
    assert eng.tuning()["resident_per_thread"] == 1
    eng.set_flags(extra=np.zeros(b.nb, np.uint8))
    assert eng.mode == MODE_RESIDENT and eng.tuning()["resident_per_thread"] == 2
    ref, orc = _pair(p, b, mode=MODE_RESIDENT, precision="f32")
    assert ref.tuning()["resident_per_thread"] == 2
    seeds = ref.get_rng().copy()
    eng.run(60)
    ref.run(60)
    np.testing.assert_array_equal(eng.get_state(), ref.get_state())
    np.testing.assert_array_equal(eng.get_rng(), ref.get_rng())
    assert orc.run(0, 60, seeds=seeds) == 0
    np.testing.assert_array_equal(ref.get_rng(), seeds)
    dx = np.abs(ref.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("resident after set_flags max |dx|", dx, flush=True)
    assert dx < 1e-3
    eng.close()
    ref.close()


@pytest.mark.parametrize("chunk", [0, TEA_SYMMETRIC])
def test_tea_drift_against_fp64_exact_em(chunk):
    print(f"case TEA drift vs FP64 exact EM (312 beads), tea_chunk={chunk}", flush=True)
    from mt_amd import params as P
    from oracle import em_exact
    # dt = 1000 ps: the zero-noise step is ~10 nm, resolvable in float32 positions
    p = P.derive(dict(P.TEA, tea_epsilon_freq=1, dt=1000.0))
    b = c1_batch()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    if chunk:
        eng.set_tuning(tea_chunk=chunk)
    orc = Oracle(p, b.copy())
    orc.rebuild()
    F = orc.forces()
    eng.set_noise(np.zeros((1, b.nb, 6), np.float32))
    c0 = b.coords.astype(np.float64)
    eng.run(1)
    g = eng.get_state().astype(np.float64)
    mob = (b.fixed == 0) & (b.extra == 0)
    exact = em_exact.drift(c0[:, :3], F[:, :3], p.tea_a, p.dt, p.gammaR)
    d_gpu = g[:, :3] - c0[:, :3]
    err = np.abs(d_gpu[mob] - exact[mob]).max() / np.abs(exact[mob]).max()
    print("TEA drift relative error vs FP64", err, flush=True)
    assert err < 1e-4
    eng.close()


@pytest.mark.parametrize("chunk", [0, TEA_SYMMETRIC])
@pytest.mark.parametrize("periodic", [False, True])
def test_tea_batch_with_extra_beads_matches_oracle(periodic, chunk):
    """Extra (reservoir) beads take no part in TEA: excluded from epsilon and
    from every pair sum, also under the minimum image of a periodic box."""
    print(f"case TEA uneven batch with extra beads, periodic={periodic}, tea_chunk={chunk}", flush=True)
    from mt_amd import params as P
    p = P.derive(dict(P.TEA, tea_epsilon_freq=3))
    b = uneven_batch()
    # mark the free dimers of trajectory 1 as reservoir (extra) beads
    s, e = int(b.off[1]), int(b.off[2])
    n_mt = 13 * 2 * 5
    b.extra[s + n_mt:e] = 1
    if periodic:
        b.box = np.array([130.0, 100.0, 180.0], np.float32)   # each larger than its system
    eng, orc = _pair(p, b, precision="f64")
    # the one-sided chunks run beside the two-sided epsilon, the symmetric pair
    # blocks beside the one-pass epsilon (the default)
    if chunk:
        eng.set_tuning(tea_chunk=chunk)
    else:
        eng.set_tuning(tea_eps=TEA_EPS_TWOSIDED)
    rng = np.random.default_rng(8)
    nz = noise(rng, 6, b.nb)
    eng.set_noise(nz)
    eng.run(6)
    assert orc.run(0, 6, noise=nz) == 0
    eps, beta = eng.tea_stats()
    np.testing.assert_allclose(eps, orc.eps_t, rtol=1e-5)
    np.testing.assert_allclose(beta, orc.beta_t, rtol=1e-5)
    g = eng.get_state()
    assert np.abs(g[:, :3] - orc.b.coords[:, :3]).max() < 1e-4
    eng.close()


@pytest.mark.parametrize("method", [TEA_EPS_ONEPASS, TEA_EPS_TWOSIDED])
@pytest.mark.parametrize("periodic", [False, True])
def test_tea_eps_over_several_bead_groups(periodic, method):
    """Epsilon and ci of trajectories spanning several 1024-bead groups of the
    one-pass epsilon kernel (1040 beads: a group and 16 beads; 2600 beads:
    three groups), against the oracle, and the steps they drive; the same
    with the two-sided epsilon kernel."""
    print(f"case TEA eps over 1040 + 2600 beads, periodic={periodic}, tea_eps={method}", flush=True)
    from mt_amd import lattice, params as P, topology
    p = P.derive(dict(P.TEA, tea_epsilon_freq=2))
    rng = np.random.default_rng(12)
    atoms = [lattice.microtubule(40), lattice.microtubule(100)]
    b = topology.make_batch(atoms, [lattice.coords7(a, 0.05, rng) for a in atoms])
    if periodic:
        b.box = np.array([400.0, 900.0], np.float32)   # each longer than its tube (320, 800 nm)
    eng, orc = _pair(p, b, precision="f64")
    assert eng.tuning()["tea_eps"] == TEA_EPS_ONEPASS
    eng.set_tuning(tea_eps=method)
    nz = noise(rng, 3, b.nb)
    eng.set_noise(nz)
    eng.run(3)
    assert orc.run(0, 3, noise=nz) == 0
    eps, beta = eng.tea_stats()
    np.testing.assert_allclose(eps, orc.eps_t, rtol=1e-5)
    np.testing.assert_allclose(beta, orc.beta_t, rtol=1e-5)
    dx = np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max()
    print("max |dx|", dx, flush=True)
    assert dx < 1e-4
    eng.close()


def _small_uneven_periodic(seed=1):
    from mt_amd import workloads
    return workloads.uneven(3, dimers=(20, 30), free=(200, 500), conc=(30.0, 50.0), seed=seed)


def test_cell_lists_match_all_pairs_periodic():
    print("case cell lists vs all-pairs, periodic uneven batch", flush=True)
    from mt_amd.engine import LISTS_CELLS
    p = base_params()
    b = _small_uneven_periodic()
    eng = Engine(p, b.copy(), mode=MODE_STREAMED, list_method=LISTS_CELLS)
    orc = Oracle(p, b.copy())
    eng.rebuild()
    assert orc.rebuild() == 0
    _lists_equal(eng, orc)
    f = eng.forces()
    fo = orc.forces()
    assert rel_err(f, fo, XYZ) < 1e-5 and rel_err(f, fo, ANG) < 1e-5
    eng.close()


def test_cell_lists_long_open_tube():
    print("case cell lists (auto) on a 2600-bead open tube", flush=True)
    p = base_params()
    b = c1_batch(ntraj=2, n_dimers=100, jitter=0.1)
    eng = Engine(p, b.copy(), mode=MODE_STREAMED)
    orc = Oracle(p, b.copy())
    eng.rebuild()
    assert orc.rebuild() == 0
    _lists_equal(eng, orc)
    eng.close()


def test_lists_tiled_rebuild_for_trajectories_beyond_lds():
    """A 1950-bead tube is all-pairs (not cells) but too large for the
    one-workgroup-per-trajectory rebuild: the j-tiled kernel builds its lists,
    next to a 312-bead tube rebuilt in the same launch."""
    print("case tiled all-pairs rebuild, 1950 + 312 beads", flush=True)
    from mt_amd import lattice, topology
    from mt_amd.engine import LISTS_ALLPAIRS
    p = base_params()
    rng = np.random.default_rng(8)
    atoms = [lattice.microtubule(75), lattice.microtubule(12)]
    b = topology.make_batch(atoms, [lattice.coords7(a, 0.1, rng) for a in atoms])
    eng = Engine(p, b.copy(), mode=MODE_STREAMED, list_method=LISTS_ALLPAIRS)
    orc = Oracle(p, b.copy())
    eng.rebuild()
    assert orc.rebuild() == 0
    _lists_equal(eng, orc)
    eng.close()


@pytest.mark.parametrize("assembly", [False, True])
def test_lists_dense_cluster_beyond_candidate_buffer(assembly):
    """120 free beads in a 20 nm cube, next to a 312-bead tube: a bead in the
    cube's interior has about 30 others within the 8 nm list radius, more
    than the 24 candidates a lane of the one-workgroup-per-trajectory rebuild
    keeps, and takes its index-order scan; edge beads keep theirs and sort
    them.  Both give the oracle's lists, entry for entry (bonded lists too in
    assembly mode)."""
    print(f"case dense cluster lists, assembly {assembly}", flush=True)
    from mt_amd import lattice, topology
    p = base_params(is_assembly=assembly)
    rng = np.random.default_rng(11)
    # chain B: free tubulin of the system, not reservoir beads (chain X), which get no lists
    atoms = [lattice.free_dimers(60, 20.0, rng, min_sep=2.0, chain="B"), lattice.microtubule(12)]
    b = topology.make_batch(atoms, [lattice.coords7(a, 0.1, rng) for a in atoms], is_assembly=assembly)
    assert not b.extra.any()
    xyz = b.coords[:120, :3]
    within = ((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(-1) < 8.0 ** 2
    assert (within.sum(1) - 1 > 24).sum() > 20   # the fallback is exercised
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_STREAMED)
    orc = Oracle(p, b.copy(), max_lj=64)
    eng.rebuild()
    assert orc.rebuild() == 0
    _lists_equal(eng, orc)
    eng.close()


def test_periodic_dynamics_and_tea():
    print("case periodic plain + TEA steps vs oracle", flush=True)
    from mt_amd import params as P
    from mt_amd.engine import LISTS_CELLS
    b = _small_uneven_periodic(seed=2)
    p = base_params()
    eng, orc = Engine(p, b.copy(), mode=MODE_STREAMED, list_method=LISTS_CELLS), Oracle(p, b.copy(), precision="f32")
    nz = noise(np.random.default_rng(3), 120, b.nb)
    eng.set_noise(nz)
    eng.run(120)
    assert orc.run(0, 120, noise=nz) == 0
    assert np.abs(eng.get_state()[:, :6] - orc.b.coords[:, :6]).max() < 1e-3
    eng.close()
    pt = P.derive(dict(P.TEA, tea_epsilon_freq=2))
    eng, orc = Engine(pt, b.copy(), mode=MODE_STREAMED, list_method=LISTS_CELLS), Oracle(pt, b.copy())
    nz = noise(np.random.default_rng(4), 4, b.nb)
    eng.set_noise(nz)
    eng.run(4)
    assert orc.run(0, 4, noise=nz) == 0
    np.testing.assert_allclose(eng.tea_stats()[0], orc.eps_t, rtol=1e-5)
    assert np.abs(eng.get_state()[:, :3] - orc.b.coords[:, :3]).max() < 1e-4
    eng.close()
