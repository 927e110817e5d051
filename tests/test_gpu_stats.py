"""Long-run statistics of the engine against the FP32 CPU oracle: 16
independent 312-bead microtubules (per-trajectory HybridTaus streams), plain
Langevin at 300 K with the reference's force field, 1000 warm-up steps and
4000 sampled steps (dt 200 ps, 0.8 us).  Trajectories diverge (different
device math), so what must agree are the statistics: time-averaged energies
per term over independent trajectories, within their standard errors, and
the temperature estimate <U_harm> per dimer ~ (3/2) kB T of the harmonic
intra-dimer springs (the estimator of the reference's temp_calc tool).  With
bit-identical noise streams and overdamped dynamics the two sides also stay
close trajectory by trajectory, so the means agree far inside their errors."""
import numpy as np
import pytest

from mt_amd import lattice, params as P
from mt_amd.engine import Engine, MODE_AUTO
from oracle.oracle import Oracle, generate_seeds
from tests.helpers import c1_batch

pytestmark = pytest.mark.gpu

KB = 0.0019872041


def test_energy_statistics_match_oracle_over_long_runs():
    print("case long-run statistics: 16 x 312 beads, 5000 steps, GPU vs FP32 oracle", flush=True)
    p = P.derive(P.BASE)
    b = c1_batch(ntraj=16, seed=99)
    b.coords = lattice.wrap_angles(b.coords)
    nt, n = b.ntraj, 312
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_AUTO, rseed=4321)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f32")
    seeds = generate_seeds(4321, 2 * b.nb)
    tid = np.repeat(np.arange(nt), n)
    warm, nblocks, every = 1000, 40, 100
    eng.run(warm)
    assert orc.run(0, warm, seeds=seeds) == 0
    step = [warm]

    mobile = (b.fixed == 0) & (b.extra == 0)

    def gpu_sample(k):
        eng.run(k)
        return eng.energies()[0]

    def orc_sample(k):
        assert orc.run(step[0], k, seeds=seeds) == 0
        step[0] += k
        orc.rebuild()   # lists of this step, as the engine's energy call uses
        return orc.energies()

    def time_means(sample):
        """Per-trajectory time averages: every term over all beads, and
        U_harm over mobile beads only (the fixed ring keeps its initial
        jitter energy)."""
        tot, harm_mobile = np.zeros((nt, 7)), np.zeros(nt)
        for _ in range(nblocks):
            e = np.asarray(sample(every), np.float64)
            tot += np.stack([np.bincount(tid, weights=e[:, c], minlength=nt) for c in range(7)], 1)
            harm_mobile += np.bincount(tid, weights=np.where(mobile, e[:, 0], 0.0), minlength=nt)
        return tot / nblocks, harm_mobile / nblocks

    g, g_harm = time_means(gpu_sample)
    o, o_harm = time_means(orc_sample)
    names = ["U_harm", "U_long", "U_lat", "U_psi", "U_fi", "U_teta", "U_lj"]
    for c, name in enumerate(names):
        se = np.sqrt(g[:, c].var(ddof=1) / nt + o[:, c].var(ddof=1) / nt)
        diff = abs(g[:, c].mean() - o[:, c].mean())
        print(f"{name}: gpu {g[:, c].mean():.4f} oracle {o[:, c].mean():.4f} se {se:.4f}", flush=True)
        assert diff <= 4 * se + 1e-3 * abs(o[:, c].mean()) + 1e-6, name
    # temperature estimate from the harmonic springs of the mobile dimers: the
    # same on both sides within its standard error, and near (3/2) kB T (not
    # exactly: the attachment-point gap also moves with the bead angles, which
    # the stiff bending terms hold, so the spring is not a free 3D oscillator)
    kT = KB * p.temp
    n_mobile_dimers = int(mobile[:n].sum()) // 2
    se = np.sqrt(g_harm.var(ddof=1) / nt + o_harm.var(ddof=1) / nt)
    assert abs(g_harm.mean() - o_harm.mean()) <= 4 * se
    for e in (g_harm, o_harm):
        ratio = e.mean() / n_mobile_dimers / (1.5 * kT)
        print(f"<U_harm> per mobile dimer / (3/2 kT) = {ratio:.4f}", flush=True)
        assert abs(ratio - 1) < 0.15, ratio
    eng.close()


# ---------------------------------------------------------------------------
# Independent ensembles: the engine and the oracle on different seeds.
OVER = dict(D_lat=1.5, D_long=6.0, Temp=600.0)


def _perm_p(stat, a, b, rng, nperm=300):
    """Permutation p-value of stat(a, b), trajectories (list items) being
    exchangeable between the two ensembles under the null hypothesis."""
    obs = stat(a, b)
    pool = a + b
    hits = 0
    for _ in range(nperm):
        idx = rng.permutation(len(pool))
        pa = [pool[i] for i in idx[:len(a)]]
        pb = [pool[i] for i in idx[len(a):]]
        hits += stat(pa, pb) >= obs
    return (hits + 1) / (nperm + 1), obs


def test_independent_ensembles_bond_energies_temperature_dissociation():
    """16 GDP microtubules of 312 beads on the GPU and 16 on the FP32 oracle,
    with different HybridTaus seeds (independent noise), 5000 steps at 600 K
    with weakened lateral and longitudinal wells (D_lat 1.5, D_long 6.0) so
    that protofilaments peel within the run.  Per trajectory: per-bead
    longitudinal and lateral bond energies sampled every 250 steps after 1000
    warm-up steps, the harmonic-spring temperature estimator, and the
    dissociation count (bonded-list entries lost since step 0, from the
    pair rebuild's lists: src/compute_cuda.cu:527-674, the count that
    src/updater.cpp:154-227 turns into mt_length).  The two ensembles agree
    by trajectory-permutation tests on the KS distance of the bond-energy
    distributions and on the means (p > 1e-3 each)."""
    from scipy import stats
    print("case independent ensembles: 16 + 16 x 312 beads, 5000 steps, different seeds", flush=True)
    p = P.derive(dict(P.BASE, **OVER))
    b = c1_batch(ntraj=16, seed=99)
    b.coords = lattice.wrap_angles(b.coords)
    b.gtp[:] = 0
    nt, n = b.ntraj, 312
    mobile = (b.fixed == 0) & (b.extra == 0)
    eng = Engine(p, b.copy(), max_lj=64, mode=MODE_AUTO, rseed=4321)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f32")
    seeds = generate_seeds(777, 2 * b.nb)
    orc.rebuild()
    lat0 = orc.b.lat_count.reshape(nt, n).sum(1)
    lon0 = orc.b.lon_count.reshape(nt, n).sum(1)
    warm, steps, every = 1000, 4000, 250
    eng.run(warm)
    assert orc.run(0, warm, seeds=seeds) == 0
    ens = {k: {"ulong": [[] for _ in range(nt)], "ulat": [[] for _ in range(nt)], "uharm": np.zeros(nt)}
           for k in ("gpu", "orc")}
    mob_t = mobile.reshape(nt, n)
    for s in range(warm, warm + steps, every):
        eng.run(every)
        eg = eng.energies()[0]
        assert orc.run(s, every, seeds=seeds) == 0
        orc.rebuild()
        eo = orc.energies()
        for k, e in (("gpu", eg), ("orc", eo)):
            e = np.asarray(e, np.float64).reshape(nt, n, 7)
            for t in range(nt):
                ens[k]["ulong"][t].append(e[t, mob_t[t], 1])
                ens[k]["ulat"][t].append(e[t, mob_t[t], 2])
            ens[k]["uharm"] += (e[:, :, 0] * mob_t).sum(1) / (steps // every)
    eng.rebuild()
    L = eng.lists()
    diss = {"gpu": (lat0 - L["lat_count"].reshape(nt, n).sum(1)) + (lon0 - L["lon_count"].reshape(nt, n).sum(1)),
            "orc": (lat0 - orc.b.lat_count.reshape(nt, n).sum(1)) + (lon0 - orc.b.lon_count.reshape(nt, n).sum(1))}
    eng.close()
    rng = np.random.default_rng(5)

    def ks(a, c):
        return stats.ks_2samp(np.concatenate([np.concatenate(x) for x in a]),
                              np.concatenate([np.concatenate(x) for x in c])).statistic

    def mean_diff(a, c):
        return abs(np.mean(a) - np.mean(c))

    kT = KB * p.temp
    n_mobile_dimers = int(mob_t[0].sum()) // 2
    t_us = (warm + steps) * p.dt * 1e-6
    report = {}
    for name in ("ulong", "ulat"):
        report[f"KS {name}"] = _perm_p(ks, ens["gpu"][name], ens["orc"][name], rng)
    report["temperature"] = _perm_p(mean_diff, list(ens["gpu"]["uharm"]), list(ens["orc"]["uharm"]), rng)
    report["dissociation"] = _perm_p(mean_diff, list(diss["gpu"].astype(float)), list(diss["orc"].astype(float)), rng)
    for k in ("gpu", "orc"):
        print(f"{k}: <U_harm>/mobile dimer / (3/2 kT) = {ens[k]['uharm'].mean() / n_mobile_dimers / (1.5 * kT):.4f}; "
              f"bonds lost {diss[k].mean():.2f} +- {diss[k].std(ddof=1) / np.sqrt(nt):.2f} per trajectory, "
              f"{diss[k].mean() / t_us:.2f} per us", flush=True)
    for name, (pv, obs) in report.items():
        print(f"{name}: statistic {obs:.4g}, permutation p = {pv:.3f}", flush=True)
        assert pv > 1e-3, name
    assert diss["gpu"].mean() > 10 and diss["orc"].mean() > 10     # dissociation happens in this run
