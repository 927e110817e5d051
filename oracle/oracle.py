"""ctypes front end of the CPU oracle (oracle/maddy_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXLON, MAXLAT = 8, 16

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


class _Topo(ctypes.Structure):
    _fields_ = [("ntraj", ctypes.c_int32), ("max_lj", ctypes.c_int32), ("off", ctypes.c_void_p),
                ("harm_count", ctypes.c_void_p), ("harm", ctypes.c_void_p),
                ("lon_count", ctypes.c_void_p), ("lon", ctypes.c_void_p),
                ("lat_count", ctypes.c_void_p), ("lat", ctypes.c_void_p),
                ("lj_count", ctypes.c_void_p), ("lj", ctypes.c_void_p),
                ("fixed", ctypes.c_void_p), ("extra", ctypes.c_void_p),
                ("mon_type", ctypes.c_void_p), ("gtp", ctypes.c_void_p), ("on_tub", ctypes.c_void_p),
                ("zs", ctypes.c_void_p), ("box", ctypes.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load(precision: str):
    path = os.path.join(HERE, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.or_forces.argtypes = [P, P, _f32p, ctypes.c_int64, _f64p]
    lib.or_energies.argtypes = [P, P, _f32p, ctypes.c_int64, _f64p]
    lib.or_pairs.argtypes = [P, _f32p, ctypes.c_int64]
    lib.or_pairs.restype = ctypes.c_int64
    lib.or_lj_list.argtypes = [P, P, _f32p, ctypes.c_int64]
    lib.or_lj_list.restype = ctypes.c_int64
    lib.or_rforce.argtypes = [_u32p, _f32p]
    lib.or_taus_draws.argtypes = [_u32p, ctypes.c_int64, _u32p]
    lib.or_generate_seeds.argtypes = [ctypes.c_int32, ctypes.c_int64, _u32p]
    lib.or_integrate.argtypes = [P, P, _f32p, _f64p, ctypes.c_int64, P, P]
    lib.or_tea_eps.argtypes = [P, P, _f32p, ctypes.c_int64, _f64p, _f64p]
    lib.or_tea_beta.argtypes = [P, P, _f64p, _f64p, _f64p]
    lib.or_tea_beta.restype = ctypes.c_int
    lib.or_tea_integrate.argtypes = [P, P, _f32p, _f64p, ctypes.c_int64, _f64p, _f64p, P, P]
    lib.or_run.argtypes = [P, P, _f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, P, P, P, P, P]
    lib.or_run.restype = ctypes.c_int64
    lib.or_max_threads.restype = ctypes.c_int
    return lib


_LIBS = {}


def lib(precision="f64"):
    if precision not in _LIBS:
        _LIBS[precision] = _load(precision)
    return _LIBS[precision]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """Holds a batch (mt_amd.topology.Batch) and its parameters for oracle calls.

    The batch's list arrays are mutated in place by rebuilds, as the
    reference's device lists are."""

    def __init__(self, params, batch, max_lj=64, precision="f64"):
        self.p = params
        self.b = batch
        self.lib = lib(precision)
        nb = batch.nb
        self.max_lj = max_lj
        self.lj_count = np.zeros(nb, np.int32)
        self.lj = np.zeros((nb, max_lj), np.int32)
        for name in ("lon_count", "lon", "lat_count", "lat", "harm_count", "harm", "mon_type", "gtp", "on_tub"):
            setattr(batch, name, np.ascontiguousarray(getattr(batch, name), np.int32))
        batch.fixed = np.ascontiguousarray(batch.fixed, np.uint8)
        batch.extra = np.ascontiguousarray(batch.extra, np.uint8)
        batch.zs = np.ascontiguousarray(batch.zs, np.float32)
        if batch.box is None or len(batch.box) != batch.ntraj:
            batch.box = np.zeros(batch.ntraj, np.float32)
        batch.box = np.ascontiguousarray(batch.box, np.float32)
        batch.off = np.ascontiguousarray(batch.off, np.int64)
        self.ci = np.zeros((nb, 3))
        self.eps_t = np.zeros(batch.ntraj)
        self.beta_t = np.zeros(batch.ntraj)
        self._topo()

    def _topo(self):
        b = self.b
        self.t = _Topo(b.ntraj, self.max_lj, _ptr(b.off), _ptr(b.harm_count), _ptr(b.harm), _ptr(b.lon_count),
                       _ptr(b.lon), _ptr(b.lat_count), _ptr(b.lat), _ptr(self.lj_count), _ptr(self.lj),
                       _ptr(b.fixed), _ptr(b.extra), _ptr(b.mon_type), _ptr(b.gtp), _ptr(b.on_tub), _ptr(b.zs),
                       _ptr(b.box))

    @property
    def _pp(self):
        return ctypes.byref(self.p)

    @property
    def _tp(self):
        return ctypes.byref(self.t)

    def forces(self, coords=None):
        c = np.ascontiguousarray(self.b.coords if coords is None else coords, np.float32)
        out = np.zeros((len(c), 7))
        self.lib.or_forces(self._pp, self._tp, c, len(c), out)
        return out

    def energies(self, coords=None):
        c = np.ascontiguousarray(self.b.coords if coords is None else coords, np.float32)
        out = np.zeros((len(c), 7))
        self.lib.or_energies(self._pp, self._tp, c, len(c), out)
        return out

    def rebuild(self, coords=None):
        c = np.ascontiguousarray(self.b.coords if coords is None else coords, np.float32)
        ov = 0
        if self.p.lj_on:
            ov += self.lib.or_lj_list(self._pp, self._tp, c, len(c))
        if self.p.is_assembly:
            ov += self.lib.or_pairs(self._tp, c, len(c))
        return ov

    def tea_eps(self, coords=None):
        c = np.ascontiguousarray(self.b.coords if coords is None else coords, np.float32)
        ci = np.zeros((len(c), 3))
        eps = np.zeros(len(c))
        self.lib.or_tea_eps(self._pp, self._tp, c, len(c), ci, eps)
        return ci, eps

    def tea_beta(self, eps_bead):
        eps_t = np.zeros(self.b.ntraj)
        beta_t = np.zeros(self.b.ntraj)
        err = self.lib.or_tea_beta(self._pp, self._tp, np.ascontiguousarray(eps_bead, np.float64), eps_t, beta_t)
        return err, eps_t, beta_t

    def run(self, step0, nsteps, noise=None, seeds=None):
        """Advance self.b.coords in place; returns the oracle's status code."""
        c = self.b.coords
        assert c.dtype == np.float32 and c.flags["C_CONTIGUOUS"]
        if noise is not None:
            noise = np.ascontiguousarray(noise, np.float32)
            assert noise.size == 6 * self.b.nb * nsteps
        return self.lib.or_run(self._pp, self._tp, c, self.b.nb, step0, nsteps, _ptr(noise), _ptr(seeds),
                               _ptr(self.ci), _ptr(self.eps_t), _ptr(self.beta_t))


def generate_seeds(seed: int, np_: int) -> np.ndarray:
    out = np.zeros(4 * np_, np.uint32)
    lib().or_generate_seeds(seed, np_, out)
    return out


def rforce(state: np.ndarray):
    out = np.zeros(4, np.float32)
    lib().or_rforce(state, out)
    return out


def taus_draws(state: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().or_taus_draws(state, n, out)
    return out


def max_threads() -> int:
    return lib().or_max_threads()
