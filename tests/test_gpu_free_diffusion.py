"""Free beads against the integrator's closed forms.  With no bonds, LJ,
barrier, walls or HI a bead feels no force or torque and only diffuses
(src/compute_cuda.cu:943-975), so after K steps each displacement is normal
with variance
  x, y, z:   K varR^2                      varR = sqrt(2 kB T dt / gammaR)
  phi, psi:  K varTheta^2 freeze_temp / alpha
  theta:     K varTheta^2                  varTheta = sqrt(2 kB T dt / gammaTheta)
(src/preparator.cpp:211-219): the free-bead MSD 6 (kB T / gammaR) t and its
rotational analogues.  This pins the noise scales of the integrator and the
normals of the HybridTaus / Box-Muller draws, in streamed and resident modes."""
import numpy as np
import pytest

from mt_amd import lattice, params as P, topology
from mt_amd.engine import Engine, MODE_RESIDENT, MODE_STREAMED

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [MODE_STREAMED, MODE_RESIDENT])
def test_free_bead_diffusion_closed_forms(mode):
    print(f"case free-bead diffusion closed forms, mode {mode}", flush=True)
    rng = np.random.default_rng(4)
    atoms, coords = [], []
    for _ in range(4):
        a = lattice.free_dimers(512, 400.0, rng, chain="B")   # chain X would make them reservoir beads
        atoms.append(a)
        coords.append(lattice.coords7(a))
    b = topology.make_batch(atoms, coords, is_assembly=False)
    b.harm_count[:] = 0
    b.harm[:] = -1
    b.lon_count[:] = 0
    b.lat_count[:] = 0
    assert not b.extra.any() and not b.fixed.any()
    p = P.derive(dict(P.BASE, LJ_on=False, barrier=False, is_assembly=False, repulsive_walls=False))
    eng = Engine(p, b.copy(), mode=mode, rseed=77)
    assert eng.mode == mode
    K = 200
    c0 = eng.get_state().astype(np.float64)
    eng.run(K)
    d = eng.get_state().astype(np.float64) - c0
    eng.close()
    n = b.nb
    vr = K * float(p.varR) ** 2
    vt = K * float(p.varTheta) ** 2
    va = vt * float(p.freeze_temp) / float(p.alpha)
    # x, y, z pooled (3n samples, relative standard error of the variance sqrt(2 / 3n) = 1.3%)
    xyz = d[:, :3].ravel()
    print(f"var xyz / K varR^2 = {xyz.var() / vr:.4f}; phi {d[:, 3].var() / va:.4f}; "
          f"theta {d[:, 4].var() / vt:.4f}; psi {d[:, 5].var() / va:.4f}", flush=True)
    assert abs(xyz.var() / vr - 1) < 0.065
    assert abs(xyz.mean()) < 5 * np.sqrt(vr / xyz.size)
    # angles (n samples each, relative standard error 2.2%)
    for col, v in ((3, va), (4, vt), (5, va)):
        assert abs(d[:, col].var() / v - 1) < 0.11, col
        assert abs(d[:, col].mean()) < 5 * np.sqrt(v / n), col
    # the sums of K normal draws are normal: excess kurtosis 0 (standard error sqrt(24 / 3n))
    z = xyz / np.sqrt(vr)
    assert abs(np.mean(z ** 4) / np.mean(z ** 2) ** 2 - 3) < 0.25
