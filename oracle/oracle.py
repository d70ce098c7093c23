"""ctypes front end of the C oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built from ``method_oracle.c`` by ``oracle/Makefile``)
for tests/, the smoke check and bench.py's CPU baseline. The product package
(``method_amd``) never imports this module.

Arrays are numpy float64 in the reference layout ``[var][i][j][k]`` over the
ghost-padded grid (Project/CPU/Include/simData.h:10).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SRMHD = 0
SRRMHD = 1
# RKPlus family (RKPlus.h:35-515): method ids of oc_rkplus_step
RKPLUS = {"rk2b": 0, "rk3": 1, "rk4": 2, "rk4_10": 3}


class OcConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("Ng", ctypes.c_int),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("gamma", ctypes.c_double), ("sigma", ctypes.c_double), ("cp", ctypes.c_double), ("cfl", ctypes.c_double),
        ("model", ctypes.c_int), ("weno_order", ctypes.c_int),
        ("periodic", ctypes.c_int * 3), ("quirks", ctypes.c_int), ("regime", ctypes.c_int),
    ]


class OcStats(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_long), ("iterations", ctypes.c_long),
                ("failures", ctypes.c_long), ("max_iterations", ctypes.c_long)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        cp = ctypes.POINTER(OcConfig)
        sp = ctypes.POINTER(OcStats)
        L.oc_apply_bcs.argtypes = [cp, dp, ctypes.c_int]
        L.oc_prims_to_all.argtypes = [cp, dp, dp, dp]
        L.oc_set.argtypes = [cp, dp, dp, dp]
        L.oc_flux_vector.argtypes = [cp, dp, dp, dp, dp, ctypes.c_int]
        L.oc_weno3_upwind.argtypes = [ctypes.c_double] * 3
        L.oc_weno3_upwind.restype = ctypes.c_double
        L.oc_weno5_upwind.argtypes = [ctypes.c_double] * 5
        L.oc_weno5_upwind.restype = ctypes.c_double
        L.oc_flux_reconstruction.argtypes = [cp, dp, dp, dp, ctypes.c_int]
        L.oc_fvs_F.argtypes = [cp, dp, dp, dp, dp]
        L.oc_source.argtypes = [cp, dp, dp, dp, dp]
        L.oc_c2p.argtypes = [cp, dp, dp, dp, sp]
        L.oc_c2p.restype = ctypes.c_long
        L.oc_srmhd_newton.argtypes = [ctypes.c_double] * 9 + [dp]
        L.oc_srmhd_newton.restype = ctypes.c_int
        L.oc_timestep.argtypes = [cp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.oc_timestep.restype = ctypes.c_double
        L.oc_rk2_step.argtypes = [cp, dp, dp, dp, ctypes.c_double, sp]
        L.oc_rk2_step.restype = ctypes.c_long
        L.oc_ssp2_step.argtypes = [cp, dp, dp, dp, ctypes.c_double, sp]
        L.oc_ssp2_step.restype = ctypes.c_long
        L.oc_rksplit_step.argtypes = [cp, dp, dp, dp, ctypes.c_double, sp]
        L.oc_rksplit_step.restype = ctypes.c_long
        L.oc_rkplus_step.argtypes = [cp, dp, dp, dp, ctypes.c_double, ctypes.c_int, sp]
        L.oc_rkplus_step.restype = ctypes.c_long
        L.oc_srmhd_max_speed.argtypes = [ctypes.c_double] * 6
        L.oc_srmhd_max_speed.restype = ctypes.c_double
        L.oc_max_wavespeed.argtypes = [cp, dp, dp, dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


NVARS = {SRMHD: (9, 8, 13), SRRMHD: (14, 11, 17)}


@dataclass
class Case:
    """One grid + model set-up, with the reference's defaults where unset."""

    nx: int
    ny: int
    nz: int
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    zmin: float = 0.0
    zmax: float = 1.0
    Ng: int = 4
    gamma: float = 4.0 / 3.0
    sigma: float = 0.0
    cp: float = 0.1
    cfl: float = 0.4
    model: int = SRMHD
    weno_order: int = 3
    periodic: tuple = (1, 0, 1)
    quirks: int = 1
    regime: int = 0

    def cfg(self) -> OcConfig:
        c = OcConfig()
        for f in ("nx", "ny", "nz", "Ng", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax",
                  "gamma", "sigma", "cp", "cfl", "model", "weno_order", "quirks", "regime"):
            setattr(c, f, getattr(self, f))
        for a in range(3):
            c.periodic[a] = int(self.periodic[a])
        return c

    @property
    def shape(self):
        Nx = self.nx + 2 * self.Ng
        Ny = self.ny + 2 * self.Ng if self.ny > 0 else 1
        Nz = self.nz + 2 * self.Ng if self.nz > 0 else 1
        return Nx, Ny, Nz

    @property
    def nvars(self):
        return NVARS[self.model]

    def zeros(self):
        nc, np_, na = self.nvars
        s = self.shape
        return (np.zeros((nc,) + s), np.zeros((np_,) + s), np.zeros((na,) + s))


class Oracle:
    """Serial CPU restatement of the reference time step for one Case."""

    def __init__(self, case: Case):
        self.case = case
        self.c = case.cfg()
        self.L = lib()

    def set(self, prims0: np.ndarray):
        cons, prims, aux = self.case.zeros()
        prims[...] = prims0
        self.L.oc_set(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux))
        return cons, prims, aux

    def prims_to_all(self, prims0: np.ndarray):
        cons, prims, aux = self.case.zeros()
        prims[...] = prims0
        self.L.oc_prims_to_all(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux))
        return cons, prims, aux

    def apply_bcs(self, arr: np.ndarray):
        self.L.oc_apply_bcs(ctypes.byref(self.c), _p(arr), arr.shape[0])

    def flux_vector(self, cons, prims, aux, dir: int, f=None):
        if f is None:
            f = np.zeros_like(cons)
        self.L.oc_flux_vector(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), _p(f), dir)
        return f

    def flux_reconstruction(self, cons, f, dir: int):
        """FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56) of padded cons and f."""
        out = np.zeros_like(np.ascontiguousarray(cons))
        self.L.oc_flux_reconstruction(ctypes.byref(self.c), _p(np.ascontiguousarray(cons)),
                                      _p(np.ascontiguousarray(f)), _p(out), dir)
        return out

    def fnet(self, cons, prims, aux):
        out = np.zeros_like(cons)
        self.L.oc_fvs_F(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), _p(out))
        return out

    def source(self, cons, prims, aux):
        out = np.zeros_like(cons)
        self.L.oc_source(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), _p(out))
        return out

    def c2p(self, cons, prims, aux):
        st = OcStats()
        fails = self.L.oc_c2p(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), ctypes.byref(st))
        return fails, st

    def timestep(self, t: float, iters: int, end_time: float = 1e10) -> float:
        return self.L.oc_timestep(ctypes.byref(self.c), t, iters, end_time)

    def max_wavespeed(self, prims, aux) -> np.ndarray:
        """Per-axis maximum fast magnetosonic speed over interior cells (SRMHD)."""
        out = np.zeros(3)
        self.L.oc_max_wavespeed(ctypes.byref(self.c), _p(prims), _p(aux), _p(out))
        return out

    def step(self, cons, prims, aux, dt: float, integrator: str = "rk2"):
        st = OcStats()
        if integrator in RKPLUS:
            fails = self.L.oc_rkplus_step(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), dt,
                                          RKPLUS[integrator], ctypes.byref(st))
            return fails, st
        fn = {"rk2": self.L.oc_rk2_step, "ssp2": self.L.oc_ssp2_step, "rksplit": self.L.oc_rksplit_step}[integrator]
        fails = fn(ctypes.byref(self.c), _p(cons), _p(prims), _p(aux), dt, ctypes.byref(st))
        return fails, st

    def run(self, prims0: np.ndarray, steps: int, integrator: str = "rk2", end_time: float = 1e10):
        """Simulation::set then `steps` updateTime calls (simulation.cc:80-133)."""
        cons, prims, aux = self.set(prims0)
        t, dts = 0.0, []
        total = OcStats()
        for it in range(steps):
            dt = self.timestep(t, it, end_time)
            fails, st = self.step(cons, prims, aux, dt, integrator)
            if fails < 0:
                raise RuntimeError("oracle step failed")
            for f in ("cells", "iterations", "failures"):
                setattr(total, f, getattr(total, f) + getattr(st, f))
            total.max_iterations = max(total.max_iterations, st.max_iterations)
            t += dt
            dts.append(dt)
        return cons, prims, aux, dts, total


def weno3(v0, v1, v2) -> float:
    return lib().oc_weno3_upwind(v0, v1, v2)


def weno5(v0, v1, v2, v3, v4) -> float:
    return lib().oc_weno5_upwind(v0, v1, v2, v3, v4)
