"""ctypes binding of the C++ host library (libmt_host.so): the reference's
configuration, PDB, DCD and XYZ files, run setup and the whole run loop
(mt_amd/host/*.cpp).  Raises if the library has not been built."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmt_host.so")
DRIVER_PATH = os.path.join(_HERE, "mt_host")
_lib = None


class HostError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HostError(f"{LIB_PATH} is missing: build it with mt_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        c, i32, pc = ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p)
        L.mth_last_error.restype = c
        L.mth_config_query.argtypes = [c, pc, ctypes.c_int, c, c, c, c, ctypes.c_int]
        L.mth_pdb_dump.argtypes = [c, c, ctypes.c_int]
        L.mth_pdb_rewrite.argtypes = [c, c]
        L.mth_xyz_rewrite.argtypes = [c, c, c, ctypes.c_int]
        L.mth_dcd_write.argtypes = [c, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_void_p]
        L.mth_setup_load.argtypes = [c, pc, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.mth_setup_free.argtypes = [ctypes.c_void_p]
        L.mth_setup_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.mth_setup_params.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params)]
        L.mth_setup_topology.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 10
        L.mth_run.argtypes = [c, pc, ctypes.c_int, i32, i32, i32]
        _lib = L
    return _lib


def _strs(items):
    arr = (ctypes.c_char_p * max(1, len(items)))(*[s.encode() for s in items])
    return arr, len(items)


def _check(rc):
    if rc != 0:
        raise HostError(lib().mth_last_error().decode())


def config_query(path, kind, name, default=None, overrides=()):
    """One lookup with the reference's semantics; raises HostError where the
    reference would die."""
    buf = ctypes.create_string_buffer(1 << 16)
    arr, n = _strs(list(overrides))
    _check(lib().mth_config_query(path.encode(), arr, n, kind.encode(), name.encode(),
                                  None if default is None else str(default).encode(), buf, len(buf)))
    return buf.value.decode()


def pdb_dump(path):
    buf = ctypes.create_string_buffer(1 << 24)
    _check(lib().mth_pdb_dump(path.encode(), buf, len(buf)))
    return buf.value.decode().splitlines()


def pdb_rewrite(src, dst):
    _check(lib().mth_pdb_rewrite(src.encode(), dst.encode()))


def xyz_rewrite(src, dst):
    buf = ctypes.create_string_buffer(1 << 20)
    _check(lib().mth_xyz_rewrite(src.encode(), dst.encode(), buf, len(buf)))
    return buf.value.decode().splitlines()


def dcd_write(path, frames, nframes_header, first, dt, freq):
    frames = np.ascontiguousarray(frames, np.float32)     # (nframes, 3, natoms)
    _check(lib().mth_dcd_write(path.encode(), frames.shape[2], nframes_header, first, dt, freq, frames.shape[0],
                               frames.ctypes.data))


class Setup:
    """A run loaded the way the reference loads it (load_run)."""

    def __init__(self, main_conf, overrides=()):
        self._h = ctypes.c_void_p()
        arr, n = _strs(list(overrides))
        _check(lib().mth_setup_load(main_conf.encode(), arr, n, ctypes.byref(self._h)))
        ntot, ntr = ctypes.c_int32(), ctypes.c_int32()
        steps, stride = ctypes.c_int64(), ctypes.c_int64()
        lib().mth_setup_sizes(self._h, ctypes.byref(ntot), ctypes.byref(ntr), ctypes.byref(steps),
                              ctypes.byref(stride))
        self.ntot, self.ntr, self.steps, self.stride = ntot.value, ntr.value, steps.value, stride.value
        self.params = Params()
        lib().mth_setup_params(self._h, ctypes.byref(self.params))
        n, nb = self.ntot, self.ntot * self.ntr
        self.harmonic_count = np.zeros(n, np.int32)
        self.harmonic = np.zeros(n, np.int32)
        self.mon_type = np.zeros(n, np.int32)
        self.fixed = np.zeros(n, np.uint8)
        self.extra = np.zeros(nb, np.uint8)
        self.lon_count = np.zeros(nb, np.int32)
        self.lon = np.zeros((nb, 8), np.int32)
        self.lat_count = np.zeros(nb, np.int32)
        self.lat = np.zeros((nb, 16), np.int32)
        self.coords = np.zeros((nb, 7), np.float32)
        lib().mth_setup_topology(self._h, *[a.ctypes.data for a in (
            self.harmonic_count, self.harmonic, self.mon_type, self.fixed, self.extra, self.lon_count, self.lon, self.lat_count, self.lat,
            self.coords)])

    def close(self):
        if self._h:
            lib().mth_setup_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()


def run(main_conf, overrides=(), mode=0, max_lj=256, quiet=True):
    """The whole run of main_conf in the working directory (what mt_host does)."""
    arr, n = _strs(list(overrides))
    _check(lib().mth_run(main_conf.encode(), arr, n, mode, max_lj, 1 if quiet else 0))


__all__ = ["HostError", "config_query", "pdb_dump", "pdb_rewrite", "xyz_rewrite", "dcd_write", "Setup", "run",
           "DRIVER_PATH"]
