"""The C++ host driver on the GPU: a run of the fixture configuration
(tests/golden/host/run: 2 trajectories of a 13x3-dimer tube with two free
dimers, 1000 steps, output every 100) through mth_run, checked against the
engine driven directly over the C ABI with the same parameters, topology and
seed: every DCD frame equals the engine state at that step, and the printed
per-trajectory energies equal mte_compute_energies there."""
import os
import re
import shutil
import struct

import numpy as np
import pytest

from mt_amd import host
from mt_amd.engine import Engine, MODE_STREAMED
from mt_amd.topology import Batch

pytestmark = pytest.mark.gpu

RUN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "host", "run")


def read_dcd(path):
    raw = open(path, "rb").read()
    n = struct.unpack_from("<i", raw, 268)[0]
    off, frames = 276, []
    while off < len(raw):
        xyz = []
        for _ in range(3):
            size = struct.unpack_from("<i", raw, off)[0]
            assert size == 4 * n
            xyz.append(np.frombuffer(raw, np.float32, n, off + 4))
            off += 8 + 4 * n
        frames.append(np.stack(xyz, 1))
    hdr = struct.unpack_from("<i4si i i i", raw, 0)
    return hdr, frames


def batch_from_setup(s):
    n, nt = s.ntot, s.ntr
    nb = n * nt
    wrapped = s.coords.copy()
    two_pi = 2 * np.pi
    for k in (3, 4, 5):
        v = wrapped[:, k].astype(np.float64)
        wrapped[:, k] = (v - two_pi * np.trunc(v / two_pi)).astype(np.float32)
    return Batch(coords=wrapped, off=np.arange(nt + 1, dtype=np.int64) * n,
                 harm_count=np.tile(s.harmonic_count, nt), harm=np.tile(s.harmonic, nt),
                 lon_count=s.lon_count.copy(), lon=s.lon.copy(), lat_count=s.lat_count.copy(), lat=s.lat.copy(),
                 fixed=np.tile(s.fixed, nt), extra=s.extra.copy(), mon_type=np.tile(s.mon_type, nt),
                 gtp=np.ones(nb, np.int32), on_tub=np.zeros(nb, np.int32), zs=np.full(nt, 160.0, np.float32))


def test_driver_run_matches_engine(tmp_path, capfd):
    print("case host driver: fixture run, streamed mode, 1000 steps", flush=True)
    for f in os.listdir(RUN):
        shutil.copy(os.path.join(RUN, f), tmp_path / f)
    old = os.getcwd()
    os.chdir(tmp_path)
    try:
        s = host.Setup("main.conf")
        capfd.readouterr()
        host.run("main.conf", mode=MODE_STREAMED, max_lj=256, quiet=True)
        out = capfd.readouterr().out
        (_, cord, nfile, npriv, nsavc, _), frames = read_dcd("traj_0.dcd")
        _, frames1 = read_dcd("traj_1.dcd")
        _, angs = read_dcd("traj_0.dcd_ang")
        assert cord == b"CORD" and (nfile, npriv, nsavc) == (10, 1, 100)
        assert len(frames) == len(frames1) == len(angs) == 10
        assert os.path.exists("result_xyz.pdb") and os.path.exists("result_ang.pdb")
        energies = [list(map(float, l.split("\t")[1:])) for l in out.splitlines() if l.startswith("Energies[")]
        assert len(energies) == 20
        eng = Engine(s.params, batch_from_setup(s), max_lj=256, mode=MODE_STREAMED, rseed=1234567)
        for k in range(10):
            st = eng.get_state()
            _, e_traj = eng.energies()
            np.testing.assert_array_equal(frames[k], st[:s.ntot, :3])
            np.testing.assert_array_equal(frames1[k], st[s.ntot:, :3])
            np.testing.assert_array_equal(angs[k], st[:s.ntot, [3, 5, 4]])   # fi, psi, theta
            for t in range(2):
                want = np.array([float("%f" % v) for v in e_traj[t]])
                np.testing.assert_array_equal(np.array(energies[2 * k + t]), want)
            eng.run(100)
        # the result PDBs hold the last output step (900), as the reference's
        res = [l for l in open("result_xyz.pdb").read().splitlines() if l.startswith("ATOM")]
        xyz = np.array([[float(l[30:38]), float(l[38:46]), float(l[46:54])] for l in res])
        np.testing.assert_allclose(xyz, frames[-1], atol=6e-4)
        eng.close()
        s.close()
    finally:
        os.chdir(old)


R_MT, R_THRES, ANG_THRES = np.float32(8.12), np.float32(16.0), np.float32(1.0)


def _classify(xyz, theta):
    """The reference's mt_length test in float32 (src/updater.cpp:164-205)."""
    rad = np.sqrt(xyz[:, 0] * xyz[:, 0] + xyz[:, 1] * xyz[:, 1])
    c = np.cos(theta.astype(np.float32))
    on = (rad < R_MT + R_THRES) & (rad > np.float32(1.0)) & (c > np.cos(ANG_THRES))
    near = np.abs(c - np.cos(ANG_THRES)) < 1e-6   # a 1-ulp cosine difference may flip these
    return on.astype(np.int32), near


def test_mt_length_on_device_matches_host_formula():
    """mte_mt_length classifies the device state as the reference's host
    loop does, counts per trajectory, and with apply=1 sets the barrier's
    flags exactly as mte_set_flags with the same flags would."""
    print("case mte_mt_length vs the host formula, 2 x 312 beads", flush=True)
    from tests.helpers import base_params, c1_batch
    p = base_params()
    b = c1_batch(ntraj=2)
    rng = np.random.default_rng(4)
    moved = rng.choice(b.nb, 40, replace=False)
    b.coords[moved[:20], 0] += 30.0            # off the tube radially
    b.coords[moved[20:], 4] += 1.2             # tilted beyond the angle threshold
    eng = Engine(p, b.copy(), max_lj=256, mode=MODE_STREAMED)
    eng.run(50)
    st = eng.get_state()
    counts, flags = eng.mt_length(R_MT + R_THRES, 1.0, np.cos(ANG_THRES), apply=True)
    want, near = _classify(st[:, :3], st[:, 4])
    assert np.array_equal(flags[~near], want[~near])
    assert 0 < want.sum() < b.nb
    np.testing.assert_array_equal(counts, [flags[:312].sum(), flags[312:].sum()])
    ref = Engine(p, b.copy(), max_lj=256, mode=MODE_STREAMED)
    ref.run(50)
    ref.set_flags(on_tubule=flags)
    np.testing.assert_array_equal(eng.energies()[1], ref.energies()[1])
    eng.run(100)
    ref.run(100)
    np.testing.assert_array_equal(eng.get_state(), ref.get_state())
    eng.close()
    ref.close()


def test_driver_mt_len_file_matches_frames(tmp_path, capfd):
    """tubule_length on: mt_len.dat holds, at every output step, 2 r_mon / 13
    times the count of beads the reference's test puts on the tube in that
    step's DCD frame."""
    print("case host driver: tubule_length yes, mt_len.dat vs DCD frames", flush=True)
    for f in os.listdir(RUN):
        shutil.copy(os.path.join(RUN, f), tmp_path / f)
    old = os.getcwd()
    os.chdir(tmp_path)
    try:
        host.run("main.conf", overrides=("tubule_length=yes",), mode=MODE_STREAMED, max_lj=256, quiet=True)
        capfd.readouterr()
        _, frames = read_dcd("traj_0.dcd")
        _, angs = read_dcd("traj_0.dcd_ang")
        _, frames1 = read_dcd("traj_1.dcd")
        _, angs1 = read_dcd("traj_1.dcd_ang")
        rows = [l.split("\t") for l in open("mt_len.dat").read().splitlines()]
        assert [int(r[0]) for r in rows] == [100 * k for k in range(10)]
        for k, r in enumerate(rows):
            for t, (fr, an) in enumerate(((frames, angs), (frames1, angs1))):
                on, near = _classify(fr[k], an[k][:, 2])   # the ang DCD holds fi, psi, theta
                if near.any():
                    continue
                assert float(r[1 + t]) == float("%f" % (np.float32(4.0) * np.float32(on.sum()) / np.float32(13)))
    finally:
        os.chdir(old)


@pytest.mark.parametrize("events", [("tubule_length=yes", "hydrolysis=yes"),
                                    ("tubule_length=yes", "is_const_conc=yes")])
def test_checkpoint_resume_continues_exactly(tmp_path, capfd, events):
    """engine_checkpoint writes the run's state at the top of step 500, before
    that step's list rebuild and events (engine state, HybridTaus streams,
    flags, the host rand() count); engine_resume replays step 500 from it, and
    every DCD frame, mt_len.dat row and printed energy from step 500 on equals
    the uninterrupted run's, with hydrolysis or the constant-concentration
    update on."""
    print(f"case host driver: checkpoint at 500, resume to 1000, {events}", flush=True)
    over = tuple(events)
    for name in ("full", "resumed"):
        d = tmp_path / name
        d.mkdir()
        (d / "dcd").mkdir()
        for f in os.listdir(RUN):
            shutil.copy(os.path.join(RUN, f), d / f)
    old = os.getcwd()
    try:
        os.chdir(tmp_path / "full")
        capfd.readouterr()
        host.run("main.conf", overrides=over + ("engine_checkpoint=500",), mode=MODE_STREAMED, max_lj=256,
                 quiet=True)
        out_full = capfd.readouterr().out
        shutil.copy("checkpoint.bin", tmp_path / "resumed" / "checkpoint.bin")
        os.chdir(tmp_path / "resumed")
        host.run("main.conf", overrides=over + ("engine_resume=checkpoint.bin",), mode=MODE_STREAMED, max_lj=256,
                 quiet=True)
        out_res = capfd.readouterr().out
        for name in ("traj_0.dcd", "traj_1.dcd", "traj_0.dcd_ang", "traj_1.dcd_ang"):
            _, full = read_dcd(str(tmp_path / "full" / name))
            _, res = read_dcd(str(tmp_path / "resumed" / name))
            assert len(full) == 10 and len(res) == 5
            for k in range(5):
                np.testing.assert_array_equal(res[k], full[5 + k])
        rows_full = open(tmp_path / "full" / "mt_len.dat").read().splitlines()
        rows_res = open(tmp_path / "resumed" / "mt_len.dat").read().splitlines()
        assert len(rows_full) == 10 and rows_res == rows_full[5:]
        e_full = [l for l in out_full.splitlines() if l.startswith("Energies[")]
        e_res = [l for l in out_res.splitlines() if l.startswith("Energies[")]
        assert len(e_full) == 20 and e_res == e_full[10:]
    finally:
        os.chdir(old)


def test_const_conc_dynamics_do_not_depend_on_output(tmp_path, capfd):
    """With is_const_conc on, the fixture's two reservoir dimers are put back
    in the box at step 100, which is also a list-rebuild step.  The lists of
    that step are built from the positions before the update (the reference
    rebuilds at the top of the step, src/compute_cuda.cu:1141-1152), so the
    trajectory is the same whether energies are printed or not."""
    print("case host driver: is_const_conc yes, output_energy yes vs no", flush=True)
    frames = {}
    for flag in ("yes", "no"):
        d = tmp_path / flag
        d.mkdir()
        for f in os.listdir(RUN):
            shutil.copy(os.path.join(RUN, f), d / f)
        old = os.getcwd()
        try:
            os.chdir(d)
            capfd.readouterr()
            host.run("main.conf", overrides=("tubule_length=yes", "is_const_conc=yes", f"output_energy={flag}"),
                     mode=MODE_STREAMED, max_lj=256, quiet=True)
            capfd.readouterr()
            frames[flag] = [read_dcd(f"traj_{t}.dcd")[1] for t in range(2)]
        finally:
            os.chdir(old)
    for t in range(2):
        # the reservoir beads moved at step 100: the update fired
        assert np.abs(frames["yes"][t][1][-4:] - frames["yes"][t][0][-4:]).max() > 1.0
        for k in range(10):
            np.testing.assert_array_equal(frames["yes"][t][k], frames["no"][t][k])
