"""Shared fixtures-as-functions for the parity tests."""
from __future__ import annotations

import numpy as np

from mt_amd import lattice, params, topology

XYZ = [0, 1, 2]
ANG = [3, 4, 5]


def rel_err(a, b, cols):
    """max |a - b| over the columns, relative to max |b| over the same columns."""
    a = np.asarray(a, np.float64)[:, cols]
    b = np.asarray(b, np.float64)[:, cols]
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / max(scale, 1e-30))


def c1_batch(ntraj=1, jitter=0.05, seed=1234567, n_dimers=12, cfg=None):
    """Config-1 style: 13 PF x n_dimers dimer microtubule(s), jittered."""
    at = lattice.microtubule(n_dimers)
    return topology.replicate(at, ntraj, jitter, seed)


def uneven_batch(seed=7, jitter=0.05, lengths=(12, 5, 20), n_free=(0, 6, 3)):
    """Trajectories of different lengths; some carry free dimers placed
    around the tube (not periodic)."""
    rng = np.random.default_rng(seed)
    atoms, coords = [], []
    for nd, nf in zip(lengths, n_free):
        a = lattice.microtubule(nd)
        if nf:
            box = 4 * nd + 40.0
            fd = lattice.free_dimers(nf, box, rng, first_id=a.n + 1, avoid=a.xyz)
            fd.xyz[:, 0:2] += 30.0  # keep them off the lattice
            a = lattice.concat(a, fd)
        atoms.append(a)
        coords.append(lattice.coords7(a, jitter, rng))
    return topology.make_batch(atoms, coords)


def base_params(**over):
    return params.derive(dict(params.BASE, **over))


def noise(rng, nsteps, nb):
    return rng.standard_normal((nsteps, nb, 6)).astype(np.float32)


# Force-field branches the default workload leaves idle: GDP and GTP bending
# offsets both non-zero (src/compute_cuda.cu:161-175, :282-283), psi and fi
# offsets, seam weakening (:447-452), repulsive walls on the cylinder and at
# both ends (:497-517).
BRANCH = dict(theta0_gdp=0.2, theta0_gtp=-0.07, psi0=0.05, fi0=-0.03, seam_coeff=2.0, repulsive_walls=True,
              rep_r=8.0, rep_eps=0.5, rep_leftborder=2.0)


def branch_params(**over):
    return base_params(**dict(BRANCH, **over))


def branch_batch(ntraj=2, jitter=0.2, seed=31, n_dimers=12, wall_z=40.0, ang_jitter=0.03):
    """c1 geometry with per-bead GTP/GDP and on/off-tubule flags drawn at
    random (so every barrier and theta0 selection case occurs), angle jitter,
    and an end wall at z = wall_z + rep_leftborder."""
    b = c1_batch(ntraj=ntraj, jitter=jitter, seed=seed, n_dimers=n_dimers)
    rng = np.random.default_rng(seed + 1)
    b.coords[:, 3:6] += rng.normal(0.0, ang_jitter, (b.nb, 3)).astype(np.float32)
    b.gtp = rng.integers(0, 2, b.nb).astype(np.int32)
    b.on_tub = rng.integers(0, 2, b.nb).astype(np.int32)
    b.zs = np.full(b.ntraj, wall_z, np.float32)
    return b
