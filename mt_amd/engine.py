"""Python front end of libmt_engine.so (include/mt_engine.h) via ctypes.

The engine is the product path: it fails loudly if the HIP library is
missing or cannot be loaded; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import Params
from .topology import Batch, MAXLAT, MAXLON

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmt_engine.so")

MTE_OK = 0
MODE_AUTO, MODE_STREAMED, MODE_RESIDENT = 0, 1, 2
LISTS_AUTO, LISTS_ALLPAIRS, LISTS_CELLS = 0, 1, 2
TEA_SYMMETRIC = 1   # tea_chunk: 64-bead chunks, each pair block evaluated once for both beads
TEA_MFMA = 2        # tea_chunk: pair sums on the matrix cores (open boundaries)
FORCES_LDS = 1      # forces: the trajectory staged in LDS (TEA / MLP-HI force pass)
FORCES_GLOBAL = 2   # forces: neighbours gathered from global memory, XCD-aware block order
TEA_EPS_ONEPASS, TEA_EPS_TWOSIDED = 1, 2   # tea_eps: epsilon terms once per pair, or once per bead of it
ERRORS = {1: "MTE_ERR_ARG", 2: "MTE_ERR_HIP", 3: "MTE_ERR_OVERFLOW", 4: "MTE_ERR_TEA", 5: "MTE_ERR_STATE",
          6: "MTE_ERR_NOISE"}


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Topology(ctypes.Structure):
    _fields_ = [("ntraj", ctypes.c_int32), ("traj_offsets", ctypes.c_void_p),
                ("harmonic_count", ctypes.c_void_p), ("harmonic", ctypes.c_void_p),
                ("longitudinal_count", ctypes.c_void_p), ("longitudinal", ctypes.c_void_p),
                ("lateral_count", ctypes.c_void_p), ("lateral", ctypes.c_void_p),
                ("fixed", ctypes.c_void_p), ("extra", ctypes.c_void_p),
                ("mon_type", ctypes.c_void_p), ("gtp", ctypes.c_void_p), ("on_tubule", ctypes.c_void_p),
                ("wall_z", ctypes.c_void_p), ("box", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [("max_lj", ctypes.c_int32), ("mode", ctypes.c_int32), ("rseed", ctypes.c_int32),
                ("device", ctypes.c_int32), ("list_method", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    """mte_tuning: launch shapes chosen by mte_autotune (0 = keep)."""
    _fields_ = [("resident_per_thread", ctypes.c_int32), ("stream_block", ctypes.c_int32),
                ("tea_chunk", ctypes.c_int32), ("cell_stencil", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("tea_eps", ctypes.c_int32), ("forces", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# Every symbol include/mt_engine.h declares, with its ctypes signature.
_P = ctypes.c_void_p
SIGNATURES = {
    "mte_create": ([_P, _P, _P, _P], ctypes.c_int),
    "mte_destroy": ([_P], ctypes.c_int),
    "mte_last_error": ([_P], ctypes.c_char_p),
    "mte_num_beads": ([_P, _P, _P], ctypes.c_int),
    "mte_set_state": ([_P, _P], ctypes.c_int),
    "mte_get_state": ([_P, _P], ctypes.c_int),
    "mte_set_flags": ([_P, _P, _P, _P], ctypes.c_int),
    "mte_mt_length": ([_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int32, _P, _P], ctypes.c_int),
    "mte_set_step": ([_P, ctypes.c_int64], ctypes.c_int),
    "mte_get_step": ([_P, _P], ctypes.c_int),
    "mte_run": ([_P, ctypes.c_int64], ctypes.c_int),
    "mte_compute_forces": ([_P, _P], ctypes.c_int),
    "mte_compute_energies": ([_P, _P, _P], ctypes.c_int),
    "mte_rebuild_lists": ([_P], ctypes.c_int),
    "mte_get_lists": ([_P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "mte_set_noise": ([_P, _P, ctypes.c_int64], ctypes.c_int),
    "mte_get_rng": ([_P, _P], ctypes.c_int),
    "mte_set_rng": ([_P, _P], ctypes.c_int),
    "mte_tea_stats": ([_P, _P, _P], ctypes.c_int),
    "mte_tea_ci": ([_P, _P], ctypes.c_int),
    "mte_set_timing": ([_P, ctypes.c_int32], ctypes.c_int),
    "mte_kernel_time": ([_P, ctypes.c_char_p, _P, _P], ctypes.c_int),
    "mte_reset_timing": ([_P], ctypes.c_int),
    "mte_get_mode": ([_P, _P], ctypes.c_int),
    "mte_get_tuning": ([_P, _P], ctypes.c_int),
    "mte_set_tuning": ([_P, _P], ctypes.c_int),
    "mte_autotune": ([_P, ctypes.c_int32, _P, ctypes.c_char_p, ctypes.c_int64], ctypes.c_int),
    "mte_mlp_set_weights": ([_P, _P, ctypes.c_int64], ctypes.c_int),
    "mte_mlp_get_weights": ([_P, _P, ctypes.c_int64], ctypes.c_int),
    "mte_mlp_local": ([_P, _P, _P], ctypes.c_int),
    "mte_mlp_forward": ([_P, _P, ctypes.c_int64, _P], ctypes.c_int),
    "mte_mlp_forward_noalloc": ([_P, _P, ctypes.c_int64, _P], ctypes.c_int),
    "mte_mlp_loss_grad": ([_P, _P, _P, ctypes.c_int64, _P, _P], ctypes.c_int),
    "mte_mlp_train": ([_P, _P, _P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, _P], ctypes.c_int),
    "mte_set_hi_mode": ([_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float], ctypes.c_int),
    "mte_mlp_status": ([_P, _P, _P, _P], ctypes.c_int),
}
MLP_NPARAM, MLP_IN, MLP_OUT = 46601, 96, 9
MLP_NOALLOC_ROWS = 8192
HI_DEFAULT, HI_MLP = 0, 1

_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libmt_engine.so; raises if it is absent (no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python mt_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """One engine instance over a CSR batch (see include/mt_engine.h)."""

    def __init__(self, params: Params, batch: Batch, max_lj: int = 64, mode: int = MODE_AUTO, rseed: int = 1234567,
                 device: int = 0, list_method: int = LISTS_AUTO):
        self.lib = load_library()
        self.params = params
        self.nb = batch.nb
        self.ntraj = batch.ntraj
        self.max_lj = max_lj
        keep = {}

        nb, nt = batch.nb, batch.ntraj
        sizes = {"off": nt + 1, "harm_count": nb, "harm": nb, "lon_count": nb, "lon": nb * MAXLON, "lat_count": nb,
                 "lat": nb * MAXLAT, "fixed": nb, "extra": nb, "mon_type": nb, "gtp": nb, "on_tub": nb, "zs": nt,
                 "box": nt}

        def arr(name, dt):
            a = np.ascontiguousarray(getattr(batch, name), dt)
            if name == "box" and a.size == 0:
                return None   # open boundaries
            if a.size != sizes[name]:
                raise ValueError(f"batch.{name} has {a.size} elements, the batch needs {sizes[name]}")
            keep[name] = a
            return _p(a)

        topo = Topology(nt, arr("off", np.int64), arr("harm_count", np.int32), arr("harm", np.int32),
                        arr("lon_count", np.int32), arr("lon", np.int32), arr("lat_count", np.int32),
                        arr("lat", np.int32), arr("fixed", np.uint8), arr("extra", np.uint8),
                        arr("mon_type", np.int32), arr("gtp", np.int32), arr("on_tub", np.int32),
                        arr("zs", np.float32), arr("box", np.float32))
        opt = Options(max_lj, mode, rseed, device, list_method)
        h = ctypes.c_void_p()
        rc = self.lib.mte_create(ctypes.byref(params), ctypes.byref(topo), ctypes.byref(opt), ctypes.byref(h))
        if rc != MTE_OK:
            raise EngineError(rc, self.lib.mte_last_error(None).decode())
        self.h = h
        self.set_state(batch.coords)

    def _chk(self, rc):
        if rc != MTE_OK:
            raise EngineError(rc, self.lib.mte_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mte_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mode(self) -> int:
        m = ctypes.c_int32()
        self._chk(self.lib.mte_get_mode(self.h, ctypes.byref(m)))
        return m.value

    def set_state(self, coords):
        c = np.ascontiguousarray(coords, np.float32)
        assert c.shape == (self.nb, 7)
        self._chk(self.lib.mte_set_state(self.h, _p(c)))

    def get_state(self):
        c = np.zeros((self.nb, 7), np.float32)
        self._chk(self.lib.mte_get_state(self.h, _p(c)))
        return c

    def set_flags(self, extra=None, gtp=None, on_tubule=None):
        def flag(name, a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            if a.size != self.nb:   # mte_set_flags reads one entry per bead
                raise ValueError(f"{name} has {a.size} elements, the batch has {self.nb} beads")
            return a
        e = flag("extra", extra, np.uint8)
        g = flag("gtp", gtp, np.int32)
        o = flag("on_tubule", on_tubule, np.int32)
        self._chk(self.lib.mte_set_flags(self.h, _p(e), _p(g), _p(o)))

    def mt_length(self, rad_hi, rad_lo, cos_min, apply=False):
        """Tubule classification on the device: (per-trajectory counts, per-bead flags)."""
        counts = np.zeros(self.ntraj, np.int32)
        flags = np.zeros(self.nb, np.int32)
        self._chk(self.lib.mte_mt_length(self.h, float(rad_hi), float(rad_lo), float(cos_min), 1 if apply else 0,
                                         _p(counts), _p(flags)))
        return counts, flags

    @property
    def step(self) -> int:
        s = ctypes.c_int64()
        self._chk(self.lib.mte_get_step(self.h, ctypes.byref(s)))
        return s.value

    @step.setter
    def step(self, v: int):
        self._chk(self.lib.mte_set_step(self.h, int(v)))

    def run(self, nsteps: int):
        self._chk(self.lib.mte_run(self.h, int(nsteps)))

    def forces(self):
        f = np.zeros((self.nb, 7), np.float32)
        self._chk(self.lib.mte_compute_forces(self.h, _p(f)))
        return f

    def energies(self):
        per_bead = np.zeros((self.nb, 7), np.float64)
        per_traj = np.zeros((self.ntraj, 7), np.float64)
        self._chk(self.lib.mte_compute_energies(self.h, _p(per_bead), _p(per_traj)))
        return per_bead, per_traj

    def rebuild(self):
        self._chk(self.lib.mte_rebuild_lists(self.h))

    def lists(self):
        nb, ml = self.nb, self.max_lj
        out = dict(lon_count=np.zeros(nb, np.int32), lon=np.zeros((nb, MAXLON), np.int32),
                   lat_count=np.zeros(nb, np.int32), lat=np.zeros((nb, MAXLAT), np.int32),
                   lj_count=np.zeros(nb, np.int32), lj=np.zeros((nb, ml), np.int32))
        self._chk(self.lib.mte_get_lists(self.h, _p(out["lon_count"]), _p(out["lon"]), _p(out["lat_count"]),
                                         _p(out["lat"]), _p(out["lj_count"]), _p(out["lj"])))
        return out

    def set_noise(self, noise):
        if noise is None:
            self._chk(self.lib.mte_set_noise(self.h, None, 0))
            return
        n = np.ascontiguousarray(noise, np.float32)
        assert n.size % (6 * self.nb) == 0
        self._chk(self.lib.mte_set_noise(self.h, _p(n), n.size // (6 * self.nb)))

    def get_rng(self):
        s = np.zeros(8 * self.nb, np.uint32)
        self._chk(self.lib.mte_get_rng(self.h, _p(s)))
        return s

    def set_rng(self, seeds):
        s = np.ascontiguousarray(seeds, np.uint32)
        assert s.size == 8 * self.nb
        self._chk(self.lib.mte_set_rng(self.h, _p(s)))

    def tea_stats(self):
        eps = np.zeros(self.ntraj)
        beta = np.zeros(self.ntraj)
        self._chk(self.lib.mte_tea_stats(self.h, _p(eps), _p(beta)))
        return eps, beta

    def tea_ci(self):
        """Per-bead row sums of squares of D_hat (B, 3) from the last TEA refresh."""
        ci = np.zeros((self.nb, 3), np.float32)
        self._chk(self.lib.mte_tea_ci(self.h, _p(ci)))
        return ci

    def tuning(self) -> dict:
        t = Tuning()
        self._chk(self.lib.mte_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        t = Tuning(**fields)
        self._chk(self.lib.mte_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, trial_steps: int = 100):
        """Time the candidate launch shapes on this batch and keep the fastest;
        returns (chosen shapes, per-trial report lines)."""
        t = Tuning()
        buf = ctypes.create_string_buffer(8192)
        self._chk(self.lib.mte_autotune(self.h, trial_steps, ctypes.byref(t), buf, len(buf)))
        return t.as_dict(), buf.value.decode().splitlines()

    def set_timing(self, on: bool):
        self._chk(self.lib.mte_set_timing(self.h, int(on)))

    def reset_timing(self):
        self._chk(self.lib.mte_reset_timing(self.h))

    def kernel_time(self, name: str):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self.lib.mte_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ---------------------------------------------------------------- MLP-HI
    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == MLP_NPARAM
        self._chk(self.lib.mte_mlp_set_weights(self.h, _p(w), w.size))

    def mlp_get_weights(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        self._chk(self.lib.mte_mlp_get_weights(self.h, _p(w), w.size))
        return w

    def mlp_local(self):
        f = np.zeros((self.nb, MLP_IN), np.float32)
        t = np.zeros((self.nb, MLP_OUT), np.float32)
        self._chk(self.lib.mte_mlp_local(self.h, _p(f), _p(t)))
        return f, t

    def mlp_forward(self, feat):
        x = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((len(x), MLP_OUT), np.float32)
        self._chk(self.lib.mte_mlp_forward(self.h, _p(x), len(x), _p(out)))
        return out

    def mlp_forward_noalloc(self, feat):
        """mte_mlp_forward_noalloc: the same inference through the device's
        fixed staging array (at most MLP_NOALLOC_ROWS rows, no allocation)."""
        x = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((len(x), MLP_OUT), np.float32)
        self._chk(self.lib.mte_mlp_forward_noalloc(self.h, _p(x), len(x), _p(out)))
        return out

    def mlp_loss_grad(self, feat, targ):
        x = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(targ, np.float32)
        loss = ctypes.c_double()
        g = np.zeros(MLP_NPARAM, np.float32)
        self._chk(self.lib.mte_mlp_loss_grad(self.h, _p(x), _p(t), len(x), ctypes.byref(loss), _p(g)))
        return loss.value, g

    def mlp_train(self, feat, targ, iters, batch, lr):
        x = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(targ, np.float32)
        losses = np.zeros(max(iters, 1), np.float64)
        self._chk(self.lib.mte_mlp_train(self.h, _p(x), _p(t), len(x), int(iters), int(batch), float(lr),
                                         _p(losses)))
        return losses[:iters]

    def set_hi_mode(self, mode, check_every=100, tolerance=0.02):
        self._chk(self.lib.mte_set_hi_mode(self.h, int(mode), int(check_every), float(tolerance)))

    def mlp_status(self):
        err = np.zeros(self.ntraj)
        fb = ctypes.c_int32()
        bad = ctypes.c_int64()
        self._chk(self.lib.mte_mlp_status(self.h, _p(err), ctypes.byref(fb), ctypes.byref(bad)))
        return err, bool(fb.value), bad.value
