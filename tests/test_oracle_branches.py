"""Pins the oracle's force-field branches with non-zero bending offsets, the
on-tubule barrier switch, seam weakening and the walls, without running the
reference.

Bending with offsets.  Reading src/compute_cuda.cu:156-177 (harmonic) and
:271-295 (longitudinal), the torque on bead i from one bond is the derivative,
with respect to i's own angle q, of a bead-side potential
    R_MON > 0 (list entry < 0):  B_q (1 - cos((q_j - q_i) - q0))
    R_MON < 0 (list entry >= 0): B_q (1 - cos((q_i - q_j) - q0))
for q = psi, fi, theta with offsets psi_0, fi_0 and theta0 (the GTP/GDP
value of bead i itself for the harmonic bond, :161,172, and of the bead with
larger z for the longitudinal bond, :282-283).  The first test checks the
oracle's torques against finite differences of that potential, in both
branches.  The second pins the energy convention of energy_kernel
(:751-757, :813-817; Appendix-A quirk Q5 of the survey): per branch the
energy's angular derivative equals the torque, either with the same offset
or with the offset negated, and the test checks which.
"""
import math

import numpy as np
import pytest

from mt_amd import params, topology
from oracle.oracle import Oracle
from tests.helpers import branch_batch, branch_params

FI, THETA, PSI = 3, 4, 5


def _decode(e):
    """(j, R_MON > 0) of a harmonic or longitudinal entry."""
    return (-e, True) if e < 0 else (e, False)


def _side_potential(b, p, c, i):
    """Bead i's bending potential per angle (fi, theta, psi), float64."""
    c = np.asarray(c, np.float64)
    U = np.zeros(3)
    bonds = []
    if b.harm_count[i]:
        j, pos = _decode(int(b.harm[i]))
        th0 = p.theta0_gtp if b.gtp[i] == 1 else p.theta0_gdp
        bonds.append((j, pos, th0))
    for k in range(int(b.lon_count[i])):
        j, pos = _decode(int(b.lon[i, k]))
        last = i if c[i, 2] > c[j, 2] else j
        th0 = p.theta0_gtp if b.gtp[last] == 1 else p.theta0_gdp
        bonds.append((j, pos, th0))
    for j, pos, th0 in bonds:
        s = 1.0 if pos else -1.0
        for n, (col, B, q0) in enumerate(((FI, p.B_fi, p.fi_0), (THETA, p.B_theta, th0), (PSI, p.B_psi, p.psi_0))):
            U[n] += B * (1.0 - math.cos(s * (c[j, col] - c[i, col]) - q0))
    return U


def _bending_torques(p, b):
    """Oracle torques from bending alone: forces minus forces with B = 0."""
    o = Oracle(p, b.copy())
    o.rebuild()
    lists = o.b
    F1 = o.forces()
    p0 = branch_params(**_zero_bending())
    o0 = Oracle(p0, lists.copy())
    o0.lj_count[:], o0.lj[:] = o.lj_count, o.lj
    return lists, F1 - o0.forces()


def _zero_bending():
    return dict(B_psi=0.0, B_fi=0.0, B_theta=0.0)


def test_bending_torques_are_side_potential_gradients():
    # small jitter: every bonded gap stays under ANGLE_CUTOFF, so every bond bends
    p = branch_params(repulsive_walls=False)
    b = branch_batch(ntraj=1, jitter=0.05, seed=41, ang_jitter=0.05)
    lists, T = _bending_torques(p, b)
    c = lists.coords.astype(np.float64)
    h = 1e-5
    worst = 0.0
    scale = np.abs(T[:, 3:6]).max()
    assert scale > 100.0
    both = {True: 0, False: 0}
    for i in range(lists.nb):
        if lists.harm_count[i]:
            both[_decode(int(lists.harm[i]))[1]] += 1
        for n, col in enumerate((FI, THETA, PSI)):
            cp, cm = c.copy(), c.copy()
            cp[i, col] += h
            cm[i, col] -= h
            g = (_side_potential(lists, p, cp, i)[n] - _side_potential(lists, p, cm, i)[n]) / (2 * h)
            worst = max(worst, abs(-g - T[i, col]) / scale)
    assert both[True] > 100 and both[False] > 100   # both R_MON branches occur
    assert worst < 1e-7


def _pair_batch(kind, gtp, ang):
    """Two beads, bonded harmonic or longitudinal: bead 0 with entry -1
    (R_MON > 0), bead 1 with entry +0 (R_MON < 0), the gap closed (bead 1 at
    z = 4 nm above bead 0 with both axes along z)."""
    c = np.zeros((2, 7), np.float32)
    c[1, 2] = 4.0
    c[:, 3:6] = np.asarray(ang, np.float32)
    z2 = np.zeros(2, np.int32)
    b = topology.Batch(coords=c, off=np.array([0, 2], np.int64), harm_count=z2.copy(), harm=np.full(2, -1, np.int32),
                       lon_count=z2.copy(), lon=np.zeros((2, 8), np.int32), lat_count=z2.copy(),
                       lat=np.zeros((2, 16), np.int32), fixed=np.zeros(2, np.uint8), extra=np.zeros(2, np.uint8),
                       mon_type=z2.copy(), gtp=np.asarray(gtp, np.int32), on_tub=np.ones(2, np.int32),
                       zs=np.zeros(1, np.float32))
    if kind == "harmonic":
        b.harm_count[:] = 1
        b.harm[:] = [-1, 0]
    else:
        b.lon_count[:] = 1
        b.lon[0, 0], b.lon[1, 0] = -1, 0
    return b


def _energy_slope(p, b, i, col, comp, h=1e-3):
    o = Oracle(p, b)
    cp, cm = b.coords.copy(), b.coords.copy()
    cp[i, col] = np.float32(cp[i, col] + h)
    cm[i, col] = np.float32(cm[i, col] - h)
    dq = float(cp[i, col]) - float(cm[i, col])
    # energy_kernel halves every term (:902-908); bead i's own term only
    return 2.0 * (o.energies(cp)[i, comp] - o.energies(cm)[i, comp]) / dq


# Which branches' energies carry the torque's own offset (+1) or the negated
# offset (-1): harmonic psi/fi use psiij, fiij and theta uses thetaji
# (:751-757); longitudinal uses psiij, fiij, thetaij (:813-817).
_CONVENTION = {
    ("harmonic", True): {FI: -1, THETA: +1, PSI: -1},
    ("harmonic", False): {FI: +1, THETA: -1, PSI: +1},
    ("longitudinal", True): {FI: -1, THETA: -1, PSI: -1},
    ("longitudinal", False): {FI: +1, THETA: +1, PSI: +1},
}
_COMP = {FI: 4, THETA: 5, PSI: 3}   # U_fi, U_teta, U_psi
_OFFSET = {FI: "fi0", PSI: "psi0"}


@pytest.mark.parametrize("kind", ["harmonic", "longitudinal"])
@pytest.mark.parametrize("gtp", [(0, 0), (1, 0), (0, 1)])
def test_bending_energy_offset_convention(kind, gtp):
    over = dict(repulsive_walls=False, LJ_on=False, barrier=False, is_assembly=False,
                theta0_gdp=0.2, theta0_gtp=-0.07, psi0=0.05, fi0=-0.03)
    p = branch_params(**over)
    p0 = branch_params(**dict(over, **_zero_bending()))
    b = _pair_batch(kind, gtp, [[0.02, -0.03, 0.04], [-0.05, 0.01, -0.02]])
    T = Oracle(p, b.copy()).forces() - Oracle(p0, b.copy()).forces()
    for i, pos in ((0, True), (1, False)):
        for col, sign in _CONVENTION[(kind, pos)].items():
            if col == THETA:
                neg = dict(over, theta0_gdp=-over["theta0_gdp"], theta0_gtp=-over["theta0_gtp"])
            else:
                neg = dict(over, **{_OFFSET[col]: -over[_OFFSET[col]]})
            same = -_energy_slope(p, b.copy(), i, col, _COMP[col])
            flipped = -_energy_slope(branch_params(**neg), b.copy(), i, col, _COMP[col])
            want, other = (same, flipped) if sign > 0 else (flipped, same)
            assert T[i, col] == pytest.approx(want, rel=2e-3, abs=1e-3), (kind, i, col)
            # the offset's sign is visible at these angles: the other convention fails
            assert abs(T[i, col] - other) > 0.05 * abs(T[i, col]) + 1.0, (kind, i, col)


def test_barrier_only_between_off_tubule_beads():
    # src/compute_cuda.cu:259,454: the barrier acts on a pair only when both
    # beads are off the tubule; with every bead on the tubule the forces equal
    # those with the barrier switched off, and with all off they differ.
    b = branch_batch(ntraj=1, jitter=0.25, seed=43)
    p_on = branch_params(repulsive_walls=False)
    p_off = branch_params(repulsive_walls=False, barrier=False)
    o = Oracle(p_on, b.copy())
    o.rebuild()
    lists = o.b

    def forces(p, on_tub):
        bb = lists.copy()
        bb.on_tub = np.full(bb.nb, on_tub, np.int32)
        oo = Oracle(p, bb)
        oo.lj_count[:], oo.lj[:] = o.lj_count, o.lj
        return oo.forces(), oo.energies()

    F_tub, E_tub = forces(p_on, 1)
    F_nob, E_nob = forces(p_off, 1)
    assert np.array_equal(F_tub, F_nob) and np.array_equal(E_tub, E_nob)
    F_free, E_free = forces(p_on, 0)
    F_free_nob, _ = forces(p_off, 0)
    assert np.abs(F_free - F_free_nob).max() > 1.0       # some gaps sit near the barrier
    assert (E_free[:, 1:3] > E_tub[:, 1:3] + 1e-9).any()  # barrier energy is positive


def test_walls_closed_forms():
    # src/compute_cuda.cu:497-517: below rep_leftborder F_z = +eps |z - lb|,
    # above zs + lb F_z = -eps |z - zt|, outside rep_r a radial push
    # -eps (rho - rep_r) along (x, y)/rho; no energy term.
    p = branch_params(LJ_on=False, barrier=False, is_assembly=False, rep_eps=0.7, rep_r=5.0, rep_leftborder=1.5)
    pts = np.array([[3.0, 4.0, -2.0], [1.0, 1.0, 3.0], [0.0, 6.0, 30.0], [0.5, 0.0, 0.5]])
    c = np.zeros((4, 7), np.float32)
    c[:, :3] = pts
    z4 = np.zeros(4, np.int32)
    b = topology.Batch(coords=c, off=np.array([0, 4], np.int64), harm_count=z4.copy(), harm=np.full(4, -1, np.int32),
                       lon_count=z4.copy(), lon=np.zeros((4, 8), np.int32), lat_count=z4.copy(),
                       lat=np.zeros((4, 16), np.int32), fixed=np.zeros(4, np.uint8), extra=np.zeros(4, np.uint8),
                       mon_type=z4.copy(), gtp=np.ones(4, np.int32), on_tub=z4.copy(),
                       zs=np.array([20.0], np.float32))
    F = Oracle(p, b).forces()
    eps, lb, zt, R = 0.7, 1.5, 21.5, 5.0
    for k, (x, y, z) in enumerate(pts):
        fz = eps * abs(z - lb) if z < lb else (-eps * abs(z - zt) if z > zt else 0.0)
        rho = math.hypot(x, y)
        fr = -eps * (rho - R) if rho > R else 0.0
        want = [fr * x / rho if rho > 0 else 0.0, fr * y / rho if rho > 0 else 0.0, fz]
        np.testing.assert_allclose(F[k, :3], want, rtol=1e-6, atol=1e-9)
