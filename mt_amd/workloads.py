"""Synthetic batches for the measured workloads.

* plain(ntraj, n_dimers): ntraj copies of one 13-PF microtubule, jittered
  N(0, 0.05 nm), open boundaries (the batched plain-Langevin and TEA runs).
* uneven(...): trajectories of different sizes, each a microtubule of
  n_d ~ U{dmin..dmax} dimers per protofilament plus n_f ~ U{fmin..fmax} free
  dimers at a concentration c ~ U[cmin, cmax] uM in a periodic cubic box of side
  (n_f / (6e-7 c))^(1/3) nm (src/updater.cpp:111), free dimers uniformly placed
  at least 8 nm apart (scripts/random_concentration_w-o_ring.py:39), CA->CB
  along +z, angles 0.  The box also holds the tube (centred, along z), so the
  side is raised to the tube length + 2 x 16 nm where needed.
"""
from __future__ import annotations

import numpy as np

from . import lattice, topology


def plain(ntraj: int, n_dimers: int, seed: int = 20240601, jitter: float = 0.05) -> topology.Batch:
    b = topology.replicate(lattice.microtubule(n_dimers), ntraj, jitter, seed)
    b.coords = lattice.wrap_angles(b.coords)
    return b


def uneven(ntraj: int, dimers=(20, 125), free=(2000, 20000), conc=(10.0, 50.0), seed: int = 4242,
           jitter: float = 0.05) -> topology.Batch:
    rng = np.random.default_rng(seed)
    atoms, coords, boxes, free_idx = [], [], [], []
    first = 0
    for _ in range(ntraj):
        nd = int(rng.integers(dimers[0], dimers[1] + 1))
        nf = int(rng.integers(free[0], free[1] + 1))
        c = float(rng.uniform(conc[0], conc[1]))
        mt = lattice.microtubule(nd)
        length = 4.0 * 2 * nd
        side = max(lattice.box_side_nm(nf, c), length + 32.0)
        mt.xyz[:, 0] += side / 2
        mt.xyz[:, 1] += side / 2
        mt.xyz[:, 2] += (side - length) / 2
        fd = lattice.free_dimers(nf, side, rng, min_sep=8.0, first_id=mt.n + 1, chain="Y", avoid=mt.xyz)
        a = lattice.concat(mt, fd)
        atoms.append(a)
        coords.append(lattice.coords7(a, jitter, rng))
        boxes.append(side)
        free_idx.append(first + mt.n + 2 * np.arange(nf))
        first += a.n
    b = topology.make_batch(atoms, coords, box=boxes)
    b.coords = lattice.wrap_angles(b.coords)
    b.free_pairs = np.concatenate(free_idx).astype(np.int64)
    return b
