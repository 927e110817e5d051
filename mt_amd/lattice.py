"""Synthetic microtubule inputs.

The geometry follows the reference's lattice generator, scripts/make_mt.py:4-66
(13 protofilaments, 13_3 lattice with a seam, monomer spacing 2*r_mon = 4 nm,
helical rise 6/13*r_mon per protofilament, radius R_MT), with the radius
R_MT = 8.12 nm of src/mt.h:23 so the compiled lateral offsets close exactly.
Free tubulin dimers follow scripts/random_concentration_w-o_ring.py and
src/updater.cpp:111-132 (CA->CB along +z, angles 0, box side from the
concentration with 6e-7 per nm^3 per uM).

Atom records carry what src/preparator.cpp:249-564 reads from a PDB: atom id,
name ("CA"/"CB"), chain letter, residue number, xyz and the three angles
(fi, psi, theta) that the reference keeps in the x, y, z columns of the
angle PDB (:307-314).
"""
from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

R_MT = 8.12
R_MON = 2.0
N_PF = 13


@dataclass
class Atoms:
    """Per-bead records in file order (one trajectory)."""
    ids: np.ndarray        # int32, PDB atom serial
    names: list            # "CA" / "CB"
    chains: np.ndarray     # uint8 chain letters
    resid: np.ndarray      # int32
    xyz: np.ndarray        # float64 (n, 3)
    ang: np.ndarray        # float64 (n, 3): fi, psi, theta (angle-PDB x, y, z)

    @property
    def n(self) -> int:
        return len(self.ids)


def microtubule(n_dimers_per_pf: int, tail_len: int = 0, theta_tail: float = 0.21) -> Atoms:
    """Ideal 13_3 microtubule with n_dimers_per_pf dimers per protofilament.

    make_mt.py's MT_len counts monomers (2 per dimer).  tail_len > 0 bends the
    top tail_len monomers outwards by theta_tail, as make_mt.py:32-39 does.
    """
    mt_len = 2 * n_dimers_per_pf
    z_step = 6.0 / 13.0 * R_MON
    alpha_step = 2.0 * math.pi / 13.0
    n = N_PF * mt_len
    ids = np.zeros(n, np.int32)
    names = []
    chains = np.zeros(n, np.uint8)
    resid = np.zeros(n, np.int32)
    xyz = np.zeros((n, 3))
    ang = np.zeros((n, 3))
    k = 0
    for c in range(N_PF):
        alpha = alpha_step * c
        for mon in range(mt_len):
            if (mt_len - mon) < tail_len:
                rad = R_MT + (tail_len - (mt_len - mon)) * 2 * R_MON * math.sin(theta_tail)
                theta = theta_tail
            else:
                rad = R_MT
                theta = 0.0
            xyz[k] = (rad * math.cos(-alpha), rad * math.sin(-alpha), 2 * R_MON * mon + z_step * c)
            ang[k] = (0.0, -alpha, theta)
            ids[k] = mon + mt_len * c + 1
            names.append("CA" if mon % 2 == 0 else "CB")
            chains[k] = ord("A") + c
            resid[k] = mon // 2 + 1
            k += 1
    return Atoms(ids, names, chains, resid, xyz, ang)


def box_side_nm(n_free_dimers: int, conc_um: float) -> float:
    """Periodic box side for n free dimers at conc_um (src/updater.cpp:111)."""
    return (n_free_dimers / (6e-7 * conc_um)) ** (1.0 / 3.0)


def free_dimers(n: int, box: float, rng: np.random.Generator, min_sep: float = 8.0,
                first_id: int = 1, chain: str = "X", avoid: np.ndarray | None = None,
                first_resid: int = 1000) -> Atoms:
    """n free dimers uniformly in [0, box)^3 with centre separation >= min_sep
    (random_concentration_w-o_ring.py:39 uses 8 nm), CA below CB along +z.
    Residue numbers start at first_resid so that `fix` (resid <= fix) never
    pins a free dimer."""
    from scipy.spatial import cKDTree
    cen = np.zeros((0, 3))
    for _ in range(50):
        need = n - len(cen)
        if need <= 0:
            break
        cand = rng.uniform(0.0, box, size=(int(need * 1.3) + 16, 3))
        pts = np.vstack([cen, cand])
        tree = cKDTree(pts % box, boxsize=box)
        drop = np.zeros(len(pts), bool)
        for i, j in tree.query_pairs(min_sep):
            # keep the earlier point (accepted ones come first)
            if not drop[i] and not drop[j]:
                drop[max(i, j)] = True
        if avoid is not None and len(avoid):
            d, _ = cKDTree(avoid % box, boxsize=box).query(pts % box, k=1)
            drop |= d < min_sep
        keep = pts[~drop]
        cen = keep[:n]
    if len(cen) < n:
        raise RuntimeError("free_dimers: box too crowded for min_sep")
    nb = 2 * n
    ids = np.arange(first_id, first_id + nb, dtype=np.int32)
    names = ["CA" if k % 2 == 0 else "CB" for k in range(nb)]
    chains = np.full(nb, ord(chain), np.uint8)
    resid = np.repeat(np.arange(first_resid, first_resid + n, dtype=np.int32), 2)
    xyz = np.zeros((nb, 3))
    xyz[0::2] = cen - np.array([0, 0, R_MON])
    xyz[1::2] = cen + np.array([0, 0, R_MON])
    ang = np.zeros((nb, 3))
    return Atoms(ids, names, chains, resid, xyz, ang)


def concat(a: Atoms, b: Atoms) -> Atoms:
    return Atoms(np.concatenate([a.ids, b.ids]), a.names + b.names, np.concatenate([a.chains, b.chains]),
                 np.concatenate([a.resid, b.resid]), np.vstack([a.xyz, b.xyz]), np.vstack([a.ang, b.ang]))


def coords7(atoms: Atoms, jitter_nm: float = 0.0, rng: np.random.Generator | None = None) -> np.ndarray:
    """Reference Coord records (x, y, z, fi, theta, psi, w) as float32 (n, 7),
    with optional N(0, jitter_nm) on xyz."""
    n = atoms.n
    c = np.zeros((n, 7), np.float64)
    c[:, 0:3] = atoms.xyz
    if jitter_nm > 0:
        assert rng is not None
        c[:, 0:3] += rng.normal(0.0, jitter_nm, size=(n, 3))
    c[:, 3] = atoms.ang[:, 0]   # fi    <- angle-PDB x
    c[:, 5] = atoms.ang[:, 1]   # psi   <- angle-PDB y
    c[:, 4] = atoms.ang[:, 2]   # theta <- angle-PDB z
    return c.astype(np.float32)


def wrap_angles(c: np.ndarray) -> np.ndarray:
    """Upload-time angle wrap toward 0 modulo 2*pi, src/compute_cuda.cu:1004-1010
    (C's (int) cast truncates toward zero)."""
    out = c.copy()
    two_pi = 2 * math.pi
    for k in (3, 4, 5):
        v = out[:, k].astype(np.float64)
        out[:, k] = (v - two_pi * np.trunc(v / two_pi)).astype(np.float32)
    return out
