"""Topology builder and CSR batching.

Restates the topology part of src/preparator.cpp:249-564 (read_PDB) and
src/preparator.cpp:717-731 (AssemblyInit) for one trajectory, and stacks
trajectories of uneven sizes into one CSR batch (global bead b of trajectory
t lives at off[t] + local index).

List entries keep the reference's encodings so lists can be compared entry by
entry with the reference's semantics:
  * harmonic partner j stored as +j when id_i > id_j, else -j; a bead with no
    partner keeps the initial fill -1 (src/preparator.cpp:339-355);
  * longitudinal entries signed by which end of i is bonded
    (src/compute_cuda.cu:587-592);
  * lateral entries signed by orientation, with ZERO = 999999 standing for
    j = 0 (src/compute_cuda.cu:646-658, src/mt.h:35).
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

from .lattice import Atoms, R_MON, R_MT

MAXLON = 8    # src/preparator.cpp:722
MAXLAT = 16   # src/preparator.cpp:721
ZERO = 999999


@dataclass
class Batch:
    """A CSR batch of trajectories: everything the hot path reads."""
    coords: np.ndarray          # (B, 7) float32, Coord order
    off: np.ndarray             # (T+1,) int64
    harm_count: np.ndarray      # (B,) int32
    harm: np.ndarray            # (B,) int32
    lon_count: np.ndarray       # (B,) int32
    lon: np.ndarray             # (B, MAXLON) int32
    lat_count: np.ndarray       # (B,) int32
    lat: np.ndarray             # (B, MAXLAT) int32
    fixed: np.ndarray           # (B,) uint8
    extra: np.ndarray           # (B,) uint8
    mon_type: np.ndarray        # (B,) int32
    gtp: np.ndarray             # (B,) int32
    on_tub: np.ndarray          # (B,) int32
    zs: np.ndarray              # (T,) float32 wall height per trajectory
    box: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # (T,) periodic side, 0 = open
    free_pairs: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))  # first bead of each free dimer

    @property
    def nb(self) -> int:
        return int(self.off[-1])

    @property
    def ntraj(self) -> int:
        return len(self.off) - 1

    def sizes(self) -> np.ndarray:
        return np.diff(self.off)

    def copy(self) -> "Batch":
        return Batch(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})


def single_topology(atoms: Atoms, fix: int = 1, is_assembly: bool = True):
    """Per-bead topology of one trajectory (read_PDB semantics)."""
    n = atoms.n
    # mon_type from name[1]: 'A' -> 0, 'B' -> 1 (src/preparator.cpp:259-265)
    mon_type = np.array([0 if nm[1] == "A" else 1 for nm in atoms.names], np.int32)
    fixed = (atoms.resid <= fix).astype(np.uint8)                  # :280-286
    extra = (atoms.chains == ord("X")).astype(np.uint8)            # :288-301
    # harmonic: same resid and chain, |i-j| < 2, i != j (:325-355)
    harm_count = np.zeros(n, np.int32)
    harm = np.full(n, -1, np.int32)
    for i in range(n):
        for j in (i - 1, i + 1):
            if 0 <= j < n and atoms.resid[i] == atoms.resid[j] and atoms.chains[i] == atoms.chains[j]:
                if harm_count[i] >= 1:
                    raise ValueError("more than one harmonic partner per bead is not supported")
                harm[i] = j if atoms.ids[i] > atoms.ids[j] else -j
                harm_count[i] += 1
    lon_count = np.zeros(n, np.int32)
    lon = np.zeros((n, MAXLON), np.int32)
    lat_count = np.zeros(n, np.int32)
    lat = np.zeros((n, MAXLAT), np.int32)
    if not is_assembly:
        _static_lists(atoms, extra, lon_count, lon, lat_count, lat)
    return dict(harm_count=harm_count, harm=harm, lon_count=lon_count, lon=lon, lat_count=lat_count, lat=lat,
                fixed=fixed, extra=extra, mon_type=mon_type)


def _rot(fi, theta, psi):
    cf, sf, ct, st, cp, sp = (math.cos(fi), math.sin(fi), math.cos(theta), math.sin(theta), math.cos(psi),
                              math.sin(psi))
    return np.array([[cp * ct, cp * st * sf - sp * cf, cp * st * cf + sp * sf],
                     [sp * ct, sp * st * sf + cp * cf, sp * st * cf - cp * sf],
                     [-st, ct * sf, ct * cf]])


def lateral_offsets():
    c, s = math.cos(2 * math.pi / 13), math.sin(2 * math.pi / 13)
    p1 = np.array([0.5 * R_MT * (c - 1), 0.5 * R_MT * s, -3 * R_MON / 13])
    p2 = np.array([0.5 * R_MT * (c - 1), -0.5 * R_MT * s, 3 * R_MON / 13])
    return p1, p2


def _static_lists(atoms, extra, lon_count, lon, lat_count, lat):
    """Non-assembly lists from the PDB (src/preparator.cpp:357-561)."""
    n = atoms.n
    for i in range(n):
        if extra[i]:
            continue
        for j in range(n):
            if (abs(int(atoms.resid[i]) - int(atoms.resid[j])) == 1 and atoms.chains[i] == atoms.chains[j]
                    and abs(int(atoms.ids[i]) - int(atoms.ids[j])) == 1):
                if lon_count[i] >= MAXLON:
                    raise ValueError("longitudinal list overflow")
                lon[i, lon_count[i]] = j if atoms.ids[i] > atoms.ids[j] else -j
                lon_count[i] += 1
    p1, p2 = lateral_offsets()
    rots = [_rot(a[0], a[2], a[1]) for a in atoms.ang]
    for i in range(n):
        if extra[i]:
            continue
        for j in range(n):
            for ind in range(2):
                a, b = (p1, p2) if ind == 0 else (p2, p1)   # ind 0: i uses p1, j uses p2
                g = atoms.xyz[i] - atoms.xyz[j] + rots[i] @ a - rots[j] @ b
                if np.linalg.norm(g) < R_MON:
                    if lat_count[i] >= MAXLAT:
                        raise ValueError("lateral list overflow")
                    lat[i, lat_count[i]] = j if ind == 1 else -j
                    lat_count[i] += 1


def make_batch(traj_atoms: list, coords: list, fix: int = 1, is_assembly: bool = True, rep_h: float = 0.0,
               box: list | None = None) -> Batch:
    """Stack per-trajectory topologies (uneven sizes allowed) into a CSR batch."""
    parts = [single_topology(a, fix, is_assembly) for a in traj_atoms]
    sizes = [a.n for a in traj_atoms]
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    nb = int(off[-1])
    return Batch(coords=np.concatenate(coords).astype(np.float32), off=off, gtp=np.ones(nb, np.int32),
                 on_tub=np.zeros(nb, np.int32), zs=np.full(len(sizes), rep_h, np.float32),
                 box=np.asarray(box if box is not None else np.zeros(len(sizes)), np.float32), **cat)


def replicate(atoms: Atoms, ntraj: int, jitter_nm: float, seed: int, fix: int = 1, is_assembly: bool = True,
              rep_h: float = 0.0) -> Batch:
    """ntraj copies of one structure, each with its own N(0, jitter) draw."""
    from .lattice import coords7
    rng = np.random.default_rng(seed)
    coords = [coords7(atoms, jitter_nm, rng) for _ in range(ntraj)]
    one = single_topology(atoms, fix, is_assembly)
    n = atoms.n
    off = np.arange(ntraj + 1, dtype=np.int64) * n
    cat = {k: np.concatenate([v] * ntraj) for k, v in one.items()}
    nb = n * ntraj
    return Batch(coords=np.concatenate(coords), off=off, gtp=np.ones(nb, np.int32), on_tub=np.zeros(nb, np.int32),
                 zs=np.full(ntraj, rep_h, np.float32), box=np.zeros(ntraj, np.float32), **cat)
