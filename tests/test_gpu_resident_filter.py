"""The resident kernel's distance-filtered LJ lists.

At each launch the resident kernel copies to LDS only the LJ entries within
LJ_CUT + 2 x 0.75 nm (+0.01) of the bead, and restores the full lists once
any bead has moved 0.75 nm from where it started the launch.  Dropped pairs
contribute exactly zero force until then, so the dynamics must equal, bit
for bit, a run whose list cutoff is short enough that nothing is filtered,
and must keep matching the unfiltered paths when a bead jumps.
"""
import numpy as np
import pytest

from mt_amd.engine import Engine, MODE_RESIDENT, MODE_STREAMED
from oracle.oracle import Oracle
from tests.helpers import base_params, c1_batch, noise

pytestmark = pytest.mark.gpu


def _run(p, b, mode, nz, max_lj=64):
    e = Engine(p, b.copy(), max_lj=max_lj, mode=mode)
    assert e.mode == mode
    e.set_noise(nz)
    e.run(nz.shape[0])
    st = e.get_state()
    e.close()
    return st


def test_filtered_lists_equal_short_cutoff_lists():
    """List cutoff 8 nm (filtered to 7.51 nm in LDS) against 7.5 nm (nothing
    filtered): the LJ sums differ only by exact zeros."""
    print("case resident filter: cutoff 8 vs 7.5, 300 steps, 2 x 1040 beads", flush=True)
    b = c1_batch(ntraj=2, n_dimers=40)
    nz = noise(np.random.default_rng(5), 300, b.nb)
    a = _run(base_params(), b, MODE_RESIDENT, nz)
    c = _run(base_params(LJPairsCutoff=7.5), b, MODE_RESIDENT, nz)
    np.testing.assert_array_equal(a, c)


def test_refill_after_a_jump_matches_streamed_and_oracle():
    """Beads pushed 2 nm in one step mid-launch: the resident run (filtered,
    then refilled) against the streamed engine (full lists) and the FP32
    oracle on the same injected noise."""
    print("case resident filter: refill after 2 nm jumps, 250 steps", flush=True)
    p = base_params()
    b = c1_batch(ntraj=2, n_dimers=40)
    nz = noise(np.random.default_rng(11), 250, b.nb)
    # step 40 and 160: twenty beads each jump about 2 nm along x (the noise is
    # scaled by varR = 0.0149 nm per unit)
    jumpers = np.random.default_rng(3).choice(b.nb, 20, replace=False)
    nz[40, jumpers, 0] += 135.0
    nz[160, jumpers, 1] -= 135.0
    r = _run(p, b, MODE_RESIDENT, nz)
    s = _run(p, b, MODE_STREAMED, nz)
    orc = Oracle(p, b.copy(), max_lj=64, precision="f32")
    assert orc.run(0, nz.shape[0], noise=nz) == 0
    o = orc.b.coords
    d_rs = np.abs(r[:, :6] - s[:, :6]).max()
    d_ro = np.abs(r[:, :3] - o[:, :3]).max()
    print("resident vs streamed max |d|", d_rs, "resident vs oracle max |dx|", d_ro, flush=True)
    assert d_ro < 1e-3
    assert d_rs < 1e-4
