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
