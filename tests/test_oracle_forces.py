"""Pins the oracle's force, energy and list code without running the
reference: lattice known answers, finite differences of the energy,
Newton's third law, invariances, and a hand-derived rotation convention.
"""
import math

import numpy as np
import pytest

from mt_amd import lattice, params, topology
from oracle.oracle import Oracle
from tests.helpers import base_params, c1_batch


def _cfg(**over):
    return params.derive(dict(params.BASE, **over))


def test_ideal_lattice_known_answers():
    # src/mt.h:44-50 offsets close the 13_3 lattice of scripts/make_mt.py exactly:
    # every bonded gap is zero, so U_long = -D_long per bond, U_lat = -D_lat per
    # lateral bond (seam_coeff 1), U_harm = 0 and the bonded forces vanish.
    p = _cfg(LJ_on=False, barrier=False)
    b = topology.replicate(lattice.microtubule(12), 1, 0.0, 1)
    o = Oracle(p, b)
    assert o.rebuild() == 0
    n = b.nb
    # 13 PF x 24 monomers: every monomer but the bottom CA / top CB of a PF has a longitudinal bond
    assert int((b.lon_count == 1).sum()) == n - 26 and int(b.lon_count.max()) == 1
    # every bead has two lateral neighbours except along the seam's ends
    assert int(b.lat_count.max()) == 2 and int((b.lat_count == 2).sum()) >= n - 13
    E = o.energies().sum(0)
    assert abs(E[0]) < 1e-6
    assert E[1] == pytest.approx(-p.D_long * b.lon_count.sum() / 2, rel=1e-6)
    assert E[2] == pytest.approx(-p.D_lat * b.lat_count.sum() / 2, rel=1e-6)
    F = o.forces()
    # residual from rounding the lattice to float32 only
    assert np.abs(F[:, :3]).max() < 5e-3 and np.abs(F[:, 3:6]).max() < 5e-3


def test_seam_weakening():
    # lateral bonds across the seam (different mon_type) use D_lat / seam_coeff
    p = _cfg(LJ_on=False, barrier=False, seam_coeff=2.0)
    b = topology.replicate(lattice.microtubule(12), 1, 0.0, 1)
    o = Oracle(p, b)
    o.rebuild()
    lat_e = o.energies()[:, 2]
    seam = 0
    for i in range(b.nb):
        for k in range(b.lat_count[i]):
            e = b.lat[i, k]
            j = abs(e)
            j = 0 if j == topology.ZERO else j
            seam += b.mon_type[i] != b.mon_type[j]
    assert seam > 0
    total = lat_e.sum() * 2
    expect = -(b.lat_count.sum() - seam) * p.D_lat - seam * p.D_lat / 2.0
    assert total == pytest.approx(expect, rel=1e-6)


def _total_energy(o, c):
    return o.energies(c).sum()


@pytest.mark.parametrize("lj", [True, False])
def test_forces_are_energy_gradients(lj):
    # With zero bending offsets every term is the gradient of the energy
    # (src/compute_cuda.cu force vs energy_kernel).  theta0_gdp != 0 breaks this
    # on purpose in the reference (energy and torque conventions differ).
    p = _cfg(theta0_gdp=0.0, LJ_on=lj, barrier=True)
    b = c1_batch(jitter=0.1, seed=3)
    b.coords[:, 3:6] += np.random.default_rng(4).normal(0, 0.02, (b.nb, 3)).astype(np.float32)
    o = Oracle(p, b)
    o.rebuild()
    F = o.forces()
    rng = np.random.default_rng(5)
    beads = rng.choice(b.nb, 12, replace=False)
    h = 1e-3
    worst = 0.0
    for i in beads:
        for q in range(6):
            cp = b.coords.copy()
            cm = b.coords.copy()
            cp[i, q] = np.float32(cp[i, q] + h)
            cm[i, q] = np.float32(cm[i, q] - h)
            dq = float(cp[i, q]) - float(cm[i, q])
            g = (_total_energy(o, cp) - _total_energy(o, cm)) / dq
            worst = max(worst, abs(-g - F[i, q]) / (np.abs(F[:, q]).max() + 1e-12))
    assert worst < 2e-4


def test_newtons_third_law_translational():
    # OutputSumForce (src/updater.cpp:45-57): pair forces cancel without walls.
    p = _cfg()
    b = c1_batch(jitter=0.1, seed=9)
    o = Oracle(p, b)
    o.rebuild()
    F = o.forces()
    s = F[:, :3].sum(0)
    assert np.abs(s).max() < 1e-9 * np.abs(F[:, :3]).sum()


def test_translation_and_z_rotation_invariance():
    p = _cfg()
    b = c1_batch(jitter=0.05, seed=13)
    o = Oracle(p, b)
    o.rebuild()
    F0 = o.forces()
    c = b.coords.astype(np.float64)
    # global translation
    t = c.copy()
    t[:, :3] += [0.5, -0.25, 1.0]
    F1 = o.forces(t.astype(np.float32))
    assert np.abs(F1 - F0).max() < 2e-3 * np.abs(F0).max()
    # rotation by beta about z together with psi += beta (make_mt.py:46-50)
    beta = 0.3
    r = c.copy()
    cb, sb = math.cos(beta), math.sin(beta)
    r[:, 0] = cb * c[:, 0] - sb * c[:, 1]
    r[:, 1] = sb * c[:, 0] + cb * c[:, 1]
    r[:, 5] += beta
    F2 = o.forces(r.astype(np.float32))
    rot = F0.copy()
    rot[:, 0] = cb * F0[:, 0] - sb * F0[:, 1]
    rot[:, 1] = sb * F0[:, 0] + cb * F0[:, 1]
    assert np.abs(F2[:, :3] - rot[:, :3]).max() < 2e-3 * np.abs(F0[:, :3]).max()
    assert np.abs(F2[:, 3:6] - F0[:, 3:6]).max() < 2e-3 * np.abs(F0[:, 3:6]).max()


def _two_bead(ci, cj, lat=None, lon=None, harm=False):
    b = topology.Batch(coords=np.array([ci, cj], np.float32), off=np.array([0, 2], np.int64),
                       harm_count=np.array([1, 1] if harm else [0, 0], np.int32),
                       harm=np.array([-1, 0] if harm else [-1, -1], np.int32),
                       lon_count=np.zeros(2, np.int32), lon=np.zeros((2, 8), np.int32),
                       lat_count=np.zeros(2, np.int32), lat=np.zeros((2, 16), np.int32),
                       fixed=np.zeros(2, np.uint8), extra=np.zeros(2, np.uint8),
                       mon_type=np.array([0, 1], np.int32), gtp=np.ones(2, np.int32), on_tub=np.zeros(2, np.int32),
                       zs=np.zeros(1, np.float32))
    if lat is not None:
        b.lat_count[0] = 1
        b.lat[0, 0] = lat
    if lon is not None:
        b.lon_count[0] = 1
        b.lon[0, 0] = lon
    return b


def test_rotation_convention_quarter_turns():
    # Known answer worked by hand from the expanded lateral gap of
    # src/compute_cuda.cu:337-352 at phi = theta = pi/2, psi = 0: the rotated
    # offset R p is (p_y, -p_z, -p_x).  Bead j (angles 0) placed at
    # r_i + R_i p2 - p1 closes the lateral gap exactly, so U_lat = -D_lat and
    # the pair force vanishes; any other Euler order leaves a gap.
    p1, p2 = topology.lateral_offsets()
    rp2 = np.array([p2[1], -p2[2], -p2[0]])
    rj = rp2 - p1
    ci = [0, 0, 0, math.pi / 2, math.pi / 2, 0, 0]      # x y z fi theta psi w
    cj = [rj[0], rj[1], rj[2], 0, 0, 0, 0]
    p = _cfg(LJ_on=False, barrier=False, is_assembly=False, seam_coeff=1.0)
    o = Oracle(p, _two_bead(ci, cj, lat=1))             # +j: i uses p2, j uses p1
    E = o.energies()
    assert E[0, 2] == pytest.approx(-p.D_lat / 2, abs=1e-5)
    assert np.abs(o.forces()[0, :3]).max() < 1e-3
    # swapped orientation (-j) pairs i's p1 with j's p2 and must not close
    o2 = Oracle(p, _two_bead(ci, cj, lat=-1))
    assert o2.energies()[0, 2] > -p.D_lat / 2 + 0.1


def test_axial_gap_known_answer():
    # Harmonic pair at theta = pi/2 for both beads: n = R e_z = (1, 0, 0)
    # (hand-derived from src/compute_cuda.cu:89-96).  With j at x = 2 r_mon,
    # d = r_j - r_i - r_mon (n_i + n_j) = 0.
    ci = [0, 0, 0, 0, math.pi / 2, 0, 0]
    cj = [4.0, 0, 0, 0, math.pi / 2, 0, 0]
    p = _cfg(LJ_on=False, barrier=False, is_assembly=False)
    o = Oracle(p, _two_bead(ci, cj, harm=True))
    E = o.energies()
    assert abs(E[0, 0]) < 1e-9
    assert np.abs(o.forces()[0, :3]).max() < 1e-3   # float32(pi/2) leaves cos(theta) ~ 4e-8


def test_radial_closed_forms():
    # Morse with zero rest length and the Gaussian barrier, src/compute_cuda.cu:16-30
    p = _cfg(LJ_on=False, barrier=True, is_assembly=False, a_barr_long=3.4, r_barr_long=0.4, w_barr_long=0.1)
    for x in (0.05, 0.3, 0.8):
        ci = [0, 0, 0, 0, 0, 0, 0]
        cj = [0, 0, 2 * topology.R_MON + x, 0, 0, 0, 0]   # j above i along z: gap x
        o = Oracle(p, _two_bead(ci, cj, lon=-1))          # -j: R_MON = +r_mon
        E = o.energies()[0, 1] * 2
        D, a = p.D_long, p.A_long
        w = p.w_barr_long
        ex = D * (1 - math.exp(-a * x)) ** 2 - D + p.a_barr_long * math.exp(-(x - p.r_barr_long) ** 2 / (2 * w * w))
        assert E == pytest.approx(ex, rel=1e-5)
        dU = 2 * a * D * (1 - math.exp(-a * x)) * math.exp(-a * x) \
            - p.a_barr_long * math.exp(-(x - p.r_barr_long) ** 2 / (2 * w * w)) * (x - p.r_barr_long) / (w * w)
        F = o.forces()
        assert F[0, 2] == pytest.approx(dU, rel=1e-4, abs=1e-6)


def test_lists_brute_force_sets_small():
    # Pair rebuild against a direct numpy enumeration of the LJ criterion
    p = _cfg()
    b = c1_batch(jitter=0.2, seed=17)
    o = Oracle(p, b)
    o.rebuild()
    x = b.coords[:, :3].astype(np.float32)
    for i in range(0, b.nb, 7):
        d = np.sqrt(((x - x[i]) ** 2).sum(1))
        expect = [j for j in range(b.nb) if j != i and d[j] < p.ljpairscutoff]
        assert list(o.lj[i, :o.lj_count[i]]) == expect


def test_periodic_image_invariance():
    # Moving any bead by a box vector leaves every force unchanged (minimum image).
    from mt_amd import workloads
    p = _cfg()
    b = workloads.uneven(2, dimers=(20, 22), free=(100, 150), conc=(40.0, 50.0), seed=5)
    o = Oracle(p, b)
    o.rebuild()
    F0 = o.forces()
    c = b.coords.copy()
    rng = np.random.default_rng(0)
    for t in range(b.ntraj):
        s, e = int(b.off[t]), int(b.off[t + 1])
        L = float(b.box[t])
        idx = rng.choice(np.arange(s, e), 50, replace=False)
        shift = rng.integers(-1, 2, size=(50, 3)) * L
        c[idx, :3] = (c[idx, :3].astype(np.float64) + shift).astype(np.float32)
    F1 = o.forces(c)
    assert np.abs(F1 - F0).max() < 5e-3 * np.abs(F0).max()


def _lj_only(jitter=0.3, **over):
    """312-bead tube with every bonded partner removed: LJ repulsion alone."""
    p = _cfg(LJ_on=True, barrier=False, is_assembly=False, **over)
    b = c1_batch(jitter=jitter)
    b.harm_count[:] = 0
    b.harm[:] = -1
    b.lon_count[:] = 0
    b.lat_count[:] = 0
    return p, b


def test_lj_regularization_cap():
    # REGULARIZATION (src/compute_cuda.cu:482-486): each LJ pair force is
    # capped at gammaR r_mon / dt (r_mon = 2 nm, src/mt.h:24).
    p_off, b = _lj_only()
    o = Oracle(p_off, b.copy(), precision="f64")
    assert o.rebuild() == 0
    F_off = o.forces()
    assert np.abs(F_off[:, 3:6]).max() == 0 and np.abs(F_off[:, :3]).max() > 0
    # at dt = 200 ps the cap (~1e4) lies far above every pair force of the
    # lattice: the switch changes nothing
    p_on, _ = _lj_only(regularization=True)
    o = Oracle(p_on, b.copy(), precision="f64")
    assert o.rebuild() == 0
    np.testing.assert_array_equal(o.forces(), F_off)
    # at dt = 1e12 ps the cap (~2e-6) lies below every pair force: each pair
    # within the 6 nm force cutoff pushes with exactly the cap along the unit
    # vector, whatever its distance
    p_cap, _ = _lj_only(regularization=True, dt=1e12)
    cap = 2.0 * float(p_cap.gammaR) / float(p_cap.dt)
    o = Oracle(p_cap, b.copy(), precision="f64")
    assert o.rebuild() == 0
    F = o.forces()[:, :3]
    x = b.coords[:, :3].astype(np.float64)
    d = x[:, None, :] - x[None, :, :]
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    near = r < 6.0
    assert near.sum() > 1000
    expect = cap * (np.where(near, 1.0 / r, 0.0)[:, :, None] * d).sum(1)
    np.testing.assert_allclose(F, expect, rtol=1e-9, atol=1e-12 * cap)
