"""Hot-path parameters and their derivation from config values.

`derive()` restates src/preparator.cpp:17-221 for the keys the hot path reads
(key names as in src/parameters.h:18-243), producing the float32 values the
reference keeps in its Parameters struct:
  gammaR = 6*pi*eta*a with a = tea_a when TEA is on, else r_mon (:211-215)
  gammaTheta = 8*pi*eta*r_mon^3 (:216)
  varR = sqrtf(2 kB T dt / gammaR), varTheta = sqrtf(2 kB T dt / gammaTheta) (:218-219)
  barrier widths: FWHM -> sigma, w / (2 sqrt(2 ln 2)) (:139, :142)
  ljsigma6 = LJSigma^6 (:161)
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

KB = 0.0019872041  # kcal/(mol K), src/mt.h:40
R_MON = 2.0


class Params(ctypes.Structure):
    """Field layout shared with include/mt_engine.h (mte_params)."""
    _fields_ = [
        ("dt", ctypes.c_float), ("temp", ctypes.c_float), ("gammaR", ctypes.c_float), ("gammaTheta", ctypes.c_float),
        ("varR", ctypes.c_float), ("varTheta", ctypes.c_float), ("alpha", ctypes.c_float),
        ("freeze_temp", ctypes.c_float),
        ("C", ctypes.c_float), ("B_psi", ctypes.c_float), ("B_fi", ctypes.c_float), ("B_theta", ctypes.c_float),
        ("psi_0", ctypes.c_float), ("fi_0", ctypes.c_float), ("theta0_gtp", ctypes.c_float),
        ("theta0_gdp", ctypes.c_float),
        ("A_long", ctypes.c_float), ("D_long", ctypes.c_float), ("A_lat", ctypes.c_float), ("D_lat", ctypes.c_float),
        ("seam_coeff", ctypes.c_float),
        ("barrier", ctypes.c_int32),
        ("a_barr_long", ctypes.c_float), ("r_barr_long", ctypes.c_float), ("w_barr_long", ctypes.c_float),
        ("a_barr_lat", ctypes.c_float), ("r_barr_lat", ctypes.c_float), ("w_barr_lat", ctypes.c_float),
        ("lj_on", ctypes.c_int32),
        ("ljscale", ctypes.c_float), ("ljsigma6", ctypes.c_float), ("ljpairscutoff", ctypes.c_float),
        ("ljpairsupdatefreq", ctypes.c_int32),
        ("is_wall", ctypes.c_int32),
        ("rep_h", ctypes.c_float), ("rep_r", ctypes.c_float), ("rep_eps", ctypes.c_float),
        ("rep_leftborder", ctypes.c_float),
        ("is_assembly", ctypes.c_int32),
        ("tea_on", ctypes.c_int32),
        ("tea_a", ctypes.c_float),
        ("tea_epsilon_freq", ctypes.c_int32), ("tea_capricious", ctypes.c_int32),
        ("tea_epsmax", ctypes.c_float),
        ("lj_regularize", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def f32(x) -> float:
    return float(np.float32(x))


def derive(cfg: dict) -> Params:
    """Build Params from a flat {key: value} config (values already typed)."""
    p = Params()
    hdi = bool(cfg.get("tea_on", False))
    tea_a = f32(cfg.get("tea_a", R_MON))
    p.dt = f32(cfg["dt"])
    p.temp = f32(cfg["Temp"])
    visc = f32(cfg["viscosity"])
    a = tea_a if hdi else R_MON
    p.gammaR = f32(6 * math.pi * visc * a)
    p.gammaTheta = f32(8 * math.pi * visc * R_MON ** 3)
    # sqrtf(2.0f*KB*Temp*dt/gammaR) evaluated in float
    two, kb = np.float32(2.0), np.float32(KB)
    p.varR = float(np.sqrt(two * kb * np.float32(p.temp) * np.float32(p.dt) / np.float32(p.gammaR)))
    p.varTheta = float(np.sqrt(two * kb * np.float32(p.temp) * np.float32(p.dt) / np.float32(p.gammaTheta)))
    p.alpha = f32(cfg["alpha"])
    p.freeze_temp = f32(cfg["freeze_temp"])
    p.C = f32(cfg["C"])
    p.B_psi, p.B_fi, p.B_theta = f32(cfg["B_psi"]), f32(cfg["B_fi"]), f32(cfg["B_theta"])
    p.psi_0, p.fi_0 = f32(cfg["psi0"]), f32(cfg["fi0"])
    p.theta0_gtp, p.theta0_gdp = f32(cfg["theta0_gtp"]), f32(cfg["theta0_gdp"])
    p.A_long, p.D_long = f32(cfg["A_long"]), f32(cfg["D_long"])
    p.A_lat, p.D_lat = f32(cfg["A_lat"]), f32(cfg["D_lat"])
    p.seam_coeff = f32(cfg["seam_coeff"])
    p.barrier = int(bool(cfg.get("barrier", False)))
    fwhm = 2 * math.sqrt(-2 * math.log(0.5))
    if p.barrier:
        p.a_barr_long, p.r_barr_long = f32(cfg["a_barr_long"]), f32(cfg["r_barr_long"])
        p.w_barr_long = f32(f32(cfg["w_barr_long"]) / fwhm)
        p.a_barr_lat, p.r_barr_lat = f32(cfg["a_barr_lat"]), f32(cfg["r_barr_lat"])
        p.w_barr_lat = f32(f32(cfg["w_barr_lat"]) / fwhm)
    p.lj_on = int(bool(cfg.get("LJ_on", False)))
    if p.lj_on:
        p.ljscale = f32(cfg["LJScale"])
        p.ljsigma6 = f32(f32(cfg["LJSigma"]) ** 6)
    p.ljpairscutoff = f32(cfg["LJPairsCutoff"])
    p.ljpairsupdatefreq = int(cfg["LJPairsUpdateFreq"])
    p.is_wall = int(bool(cfg.get("repulsive_walls", False)))
    p.rep_h, p.rep_r = f32(cfg.get("rep_h", 0)), f32(cfg.get("rep_r", 0))
    p.rep_eps, p.rep_leftborder = f32(cfg.get("rep_eps", 0)), f32(cfg.get("rep_leftborder", 0))
    p.is_assembly = int(bool(cfg.get("is_assembly", True)))
    p.tea_on = int(hdi)
    p.tea_a = tea_a
    p.tea_epsilon_freq = int(cfg.get("tea_epsilon_freq", 0)) if hdi else 1
    p.tea_capricious = int(bool(cfg.get("tea_capricious", True)))
    p.tea_epsmax = f32(cfg.get("tea_epsmax", 999.0))
    p.lj_regularize = int(bool(cfg.get("regularization", False)))
    return p


# Synthetic workloads.  Force-field values: config/morse.conf:7-11 (Morse),
# template/morse.conf:24-41 (bending, barrier), the LJ sigma of 4 nm and the
# thermostat of template/cond.conf.
BASE = dict(
    dt=200.0, Temp=300.0, viscosity=2.85e4, alpha=5.0, freeze_temp=1.0,
    C=300.0, B_psi=9125.0, B_fi=9125.0, B_theta=90.0, psi0=0.0, fi0=0.0, theta0_gtp=0.0, theta0_gdp=0.2,
    A_long=3.0, D_long=11.5, A_lat=4.0, D_lat=6.5, seam_coeff=1.0,
    barrier=True, a_barr_long=3.4, r_barr_long=0.4, w_barr_long=0.10, a_barr_lat=1.9, r_barr_lat=0.4,
    w_barr_lat=0.15,
    LJ_on=True, LJScale=0.5, LJSigma=4.0, LJPairsCutoff=8.0, LJPairsUpdateFreq=100,
    repulsive_walls=False, rep_h=160.0, rep_r=80.0, rep_eps=0.005, rep_leftborder=0.0,
    is_assembly=True, rseed=1234567, fix=1,
)

# TEA workload: water viscosity 0.89 cP = 128 kcal ps/(mol nm^3) with dt = 1 ps
# (explicit Euler is unstable at 200 ps for that viscosity), a = 2 nm.
TEA = dict(BASE, dt=1.0, viscosity=128.0, tea_on=True, tea_a=2.0, tea_epsilon_freq=100, tea_capricious=True,
           tea_epsmax=999.0)
