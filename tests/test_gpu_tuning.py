"""The autotuner (mte_autotune): it times the candidate launch shapes on the
engine's own batch, restores coordinates, RNG streams and the step counter
after every trial, and leaves results as an untuned engine would produce
them -- bit for bit where the shape does not change summation order
(resident beads per thread, streamed workgroup size, cell stencil), within
FP32 rounding where it does (TEA j-chunks regroup the pair sums)."""
import numpy as np
import pytest

from mt_amd import params as P, workloads
from mt_amd.engine import (Engine, EngineError, FORCES_GLOBAL, FORCES_LDS, MODE_RESIDENT, MODE_STREAMED,
                           TEA_SYMMETRIC)
from tests.helpers import c1_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [MODE_RESIDENT, MODE_STREAMED])
def test_autotune_plain_keeps_results(mode):
    """Plain dynamics: every resident plan and every streamed workgroup size
    is timed; the run afterwards equals the untuned engine's -- bit for bit
    when the mode stays, to FP32 summation order when it switches (the two
    paths sum the LJ terms in different orders)."""
    print(f"case autotune plain mode={mode}", flush=True)
    p = P.derive(P.BASE)
    b = c1_batch(ntraj=8, n_dimers=40)     # 1040 beads, 26 fixed: resident plans with 1, 2 or 3 beads per thread
    a = Engine(p, b.copy(), max_lj=48, mode=mode, rseed=77)
    start_mode = a.mode
    a.run(300)
    want = a.get_state()
    t = Engine(p, b.copy(), max_lj=48, mode=mode, rseed=77)
    chosen, report = t.autotune(50)
    print(chosen, report, flush=True)
    assert t.tuning() == chosen and chosen["mode"] in (MODE_RESIDENT, MODE_STREAMED)
    assert sum(l.startswith("stream_block") for l in report) == 3
    assert sum(l.startswith("resident_per_thread") for l in report) == 3   # 1, 2 and 3 beads per thread
    t.run(300)
    got = t.get_state()
    if chosen["mode"] == start_mode:
        np.testing.assert_array_equal(got, want)
    else:
        assert np.abs(got[:, :3] - want[:, :3]).max() < 1e-3
    # the tuned engine runs exactly what the chosen shapes say
    x = Engine(p, b.copy(), max_lj=48, mode=mode, rseed=77)
    x.set_tuning(**chosen)
    x.run(300)
    np.testing.assert_array_equal(x.get_state(), got)
    a.close()
    t.close()
    x.close()


def test_autotune_tea_and_cells():
    print("case autotune TEA + cell stencil (periodic uneven batch)", flush=True)
    p = P.derive(dict(P.TEA))
    b = workloads.uneven(3, dimers=(20, 30), free=(1500, 2500), conc=(30.0, 50.0), seed=5)
    a = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=9)
    a.run(120)
    want = a.get_state()
    t = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=9)
    chosen, report = t.autotune(20)
    print(chosen, report, flush=True)
    assert chosen["tea_chunk"] in (64, 128, 256, TEA_SYMMETRIC) and chosen["cell_stencil"] in (1, 2)
    assert sum(l.startswith("tea_chunk") for l in report) == 4
    assert sum(l.startswith("cell_stencil") for l in report) == 2
    t.run(120)
    got = t.get_state()
    # j-chunks regroup FP32 pair sums: the runs agree to a few float ulps of
    # the coordinates (boxes of ~450 nm), far inside what any missing pair
    # contribution would move (dt/gammaR |D F| ~ 1e-3 nm per step)
    tol = 4 * float(np.spacing(np.float32(np.abs(want[:, :3]).max())))
    print("tuned vs default, max |dx| =", np.abs(got[:, :3] - want[:, :3]).max(), "tol", tol, flush=True)
    assert np.abs(got[:, :3] - want[:, :3]).max() <= tol
    # the tuned engine runs exactly what the chosen shapes say
    x = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=9)
    x.set_tuning(**chosen)
    x.run(120)
    np.testing.assert_array_equal(x.get_state(), got)
    x.close()
    # each shape on its own (and after other shapes), against the default run
    for shapes in ([{"tea_chunk": 64}], [{"tea_chunk": 256}], [{"tea_chunk": 256}, {"tea_chunk": 64}],
                   [{"tea_chunk": TEA_SYMMETRIC}], [{"tea_chunk": TEA_SYMMETRIC}, {"tea_chunk": 128}],
                   [{"cell_stencil": 2}]):
        e = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=9)
        for shape in shapes:
            e.set_tuning(**shape)
        e.run(120)
        s = e.get_state()
        if "cell_stencil" in shapes[-1]:
            np.testing.assert_array_equal(s, want)
        else:
            assert np.abs(s[:, :3] - want[:, :3]).max() <= tol, shapes
        e.close()
    a.close()
    t.close()


def test_tea_force_kernels_agree_bit_for_bit():
    """The TEA force pass with the trajectory staged in LDS and with global
    gathers in the XCD-aware block order evaluates the same bead_force on the
    same values: trajectories agree bit for bit, at every workgroup size, and
    the autotuner times both."""
    print("case TEA force pass: LDS-staged vs global gathers", flush=True)
    p = P.derive(dict(P.TEA, tea_epsilon_freq=10))
    b = c1_batch(ntraj=5, n_dimers=20)
    ref = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=3)
    assert ref.tuning()["forces"] == FORCES_LDS
    ref.run(60)
    want = ref.get_state()
    ref.close()
    for bs in (64, 128, 256):
        e = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=3)
        e.set_tuning(forces=FORCES_GLOBAL, stream_block=bs)
        e.run(60)
        np.testing.assert_array_equal(e.get_state(), want)
        e.close()
    t = Engine(p, b.copy(), max_lj=48, mode=MODE_STREAMED, rseed=3)
    chosen, report = t.autotune(5)
    print(chosen, report, flush=True)
    assert sum(l.startswith("forces") for l in report) == 2
    assert chosen["forces"] in (FORCES_LDS, FORCES_GLOBAL)
    t.close()


def test_set_tuning_rejects_unusable_shapes():
    print("case set_tuning argument checks", flush=True)
    p = P.derive(P.BASE)
    b = c1_batch(ntraj=2, n_dimers=40)
    b.fixed[:] = 0
    e = Engine(p, b, max_lj=48, mode=MODE_RESIDENT)
    with pytest.raises(EngineError):
        e.set_tuning(stream_block=100)
    with pytest.raises(EngineError):
        e.set_tuning(resident_per_thread=1)     # 1040 mobile beads do not fit one per thread
    with pytest.raises(EngineError):
        e.set_tuning(tea_chunk=128)              # no TEA in this run
    with pytest.raises(EngineError):
        e.set_tuning(tea_eps=1)                  # no TEA in this run
    with pytest.raises(EngineError):
        e.set_tuning(forces=3)
    e.set_tuning(resident_per_thread=3)
    assert e.tuning()["resident_per_thread"] == 3
    e.close()
