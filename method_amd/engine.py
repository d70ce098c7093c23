"""Python front end of the engine's C ABI (include/method_gpu.h) via ctypes.

The engine is the GPU body behind the reference's time-step API
(Project/CPU/Include/simulation.h:84-115, timeInt.h:61-132, model.h:55-133,
flux.h:47-72). Arrays cross the boundary as numpy float64 in the reference
layout ``[var][i][j][k]`` over the ghost-padded grid (simData.h:10).

There is no CPU fallback: if ``libmethod_gpu.so`` is missing or no GPU is
present, constructing an Engine raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmethod_gpu.so")

SRMHD, SRRMHD = 0, 1
RK2, SSP2, RKSPLIT = 0, 1, 2
RK2B, RK3, RK4, RK4_10 = 3, 4, 5, 6  # RKPlus family (RKPlus.h:35-515)
EXT_NONE, EXT_REGIME = 0, 1  # model extensions (modelExtension.h, REGIME.h:81)
OUTFLOW, PERIODIC = 0, 1

EXPORTED = [
    "mgpu_last_error", "mgpu_create", "mgpu_destroy", "mgpu_sizes", "mgpu_upload_state", "mgpu_download_state",
    "mgpu_prims_to_all", "mgpu_step", "mgpu_update_time", "mgpu_get_time", "mgpu_set_time", "mgpu_fnet",
    "mgpu_flux_vector", "mgpu_source", "mgpu_flux_reconstruction", "mgpu_c2p_cells", "mgpu_c2p_cells_noalloc", "mgpu_source_cells", "mgpu_set_model_extension", "mgpu_source_extension", "mgpu_regime_fields", "mgpu_c2p", "mgpu_max_wavespeed", "mgpu_set_c2p_guess", "mgpu_set_c2p_histogram", "mgpu_get_c2p_histogram", "mgpu_timed_c2p", "mgpu_set_sweep_plan", "mgpu_set_regime_plan", "mgpu_sync", "mgpu_timed_steps",
    "mgpu_kernel_slot_name", "mgpu_mlp_create", "mgpu_mlp_destroy", "mgpu_mlp_set_data", "mgpu_mlp_train",
    "mgpu_mlp_gradients", "mgpu_mlp_get_params", "mgpu_mlp_set_params", "mgpu_mlp_predict",
    "mgpu_batch_create", "mgpu_batch_destroy", "mgpu_batch_sizes", "mgpu_batch_set_prims", "mgpu_batch_upload_state",
    "mgpu_batch_download_state", "mgpu_batch_step", "mgpu_batch_get_time", "mgpu_batch_use_graph",
    "mgpu_batch_get_tuning", "mgpu_batch_set_tuning", "mgpu_batch_autotune", "mgpu_batch_timed_steps",
    "mgpu_batch_sync",
]

GUESS_PLAIN, GUESS_LEARNED, GUESS_COLD = 0, 1, 2
MLP_PARAMS = 4866


class MgpuConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("Ng", ctypes.c_int),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double),
        ("ymax", ctypes.c_double), ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("gamma", ctypes.c_double), ("sigma", ctypes.c_double), ("cp", ctypes.c_double), ("cfl", ctypes.c_double),
        ("model", ctypes.c_int), ("weno_order", ctypes.c_int), ("integrator", ctypes.c_int),
        ("bc", ctypes.c_int * 3), ("quirks", ctypes.c_int), ("device", ctypes.c_int),
    ]


class MgpuBatchConfig(ctypes.Structure):
    _fields_ = [("Ng", ctypes.c_int), ("gamma", ctypes.c_double), ("weno_order", ctypes.c_int),
                ("bc", ctypes.c_int * 2), ("quirks", ctypes.c_int), ("device", ctypes.c_int)]


class MgpuInstance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double), ("cfl", ctypes.c_double),
                ("end_time", ctypes.c_double)]


class MgpuBatchTuning(ctypes.Structure):
    _fields_ = [("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("schedule", ctypes.c_int),
                ("blocks", ctypes.c_int), ("ms_per_step", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MgpuStats(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_longlong), ("iterations", ctypes.c_longlong),
                ("failures", ctypes.c_longlong), ("max_iterations", ctypes.c_longlong),
                ("fallbacks", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class EngineError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libmethod_gpu.so and declares the C ABI; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineError(f"HIP engine library not built: {path} (run `python -m method_amd.build`)")
    L = ctypes.CDLL(path)
    E = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    sp = ctypes.POINTER(MgpuStats)
    L.mgpu_last_error.restype = ctypes.c_char_p
    L.mgpu_create.argtypes = [ctypes.POINTER(MgpuConfig), ctypes.POINTER(E)]
    L.mgpu_destroy.argtypes = [E]
    L.mgpu_sizes.argtypes = [E, ctypes.POINTER(ctypes.c_int)] * 1 + [ctypes.POINTER(ctypes.c_int)] * 2
    L.mgpu_upload_state.argtypes = [E, dp, dp, dp]
    L.mgpu_download_state.argtypes = [E, dp, dp, dp]
    L.mgpu_prims_to_all.argtypes = [E]
    L.mgpu_step.argtypes = [E, ctypes.c_double, sp]
    L.mgpu_update_time.argtypes = [E, ctypes.c_double, sp]
    L.mgpu_get_time.argtypes = [E, dp, dp, ctypes.POINTER(ctypes.c_int)]
    L.mgpu_set_time.argtypes = [E, ctypes.c_double, ctypes.c_int]
    L.mgpu_fnet.argtypes = [E, dp]
    L.mgpu_flux_vector.argtypes = [E, ctypes.c_int, dp]
    L.mgpu_source.argtypes = [E, dp]
    L.mgpu_flux_reconstruction.argtypes = [E, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    L.mgpu_c2p_cells.argtypes = [E, ctypes.c_longlong, dp, dp, dp, sp]
    vp = ctypes.c_void_p
    L.mgpu_c2p_cells_noalloc.argtypes = [E, ctypes.c_longlong, vp, vp, vp, sp]
    L.mgpu_source_cells.argtypes = [E, ctypes.c_longlong, dp, dp, dp, dp]
    L.mgpu_set_model_extension.argtypes = [E, ctypes.c_int]
    L.mgpu_source_extension.argtypes = [E, dp]
    L.mgpu_regime_fields.argtypes = [E, dp, dp, dp, dp, dp]
    L.mgpu_c2p.argtypes = [E, sp]
    L.mgpu_max_wavespeed.argtypes = [E, dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.mgpu_set_c2p_guess.argtypes = [E, ctypes.c_int, fp, fp]
    L.mgpu_sync.argtypes = [E]
    L.mgpu_set_sweep_plan.argtypes = [E, ctypes.c_int]
    L.mgpu_set_regime_plan.argtypes = [E, ctypes.c_int]
    L.mgpu_timed_c2p.argtypes = [E, ctypes.c_int, dp, sp]
    L.mgpu_set_c2p_histogram.argtypes = [E, ctypes.c_int]
    L.mgpu_get_c2p_histogram.argtypes = [E, ctypes.POINTER(ctypes.c_longlong)]
    L.mgpu_mlp_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(E)]
    L.mgpu_mlp_destroy.argtypes = [E]
    L.mgpu_mlp_set_data.argtypes = [E, dp, dp, ctypes.c_longlong]
    L.mgpu_mlp_train.argtypes = [E, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
    L.mgpu_mlp_gradients.argtypes = [E, fp, fp, ctypes.c_int, fp, fp]
    L.mgpu_mlp_get_params.argtypes = [E, fp, fp]
    L.mgpu_mlp_set_params.argtypes = [E, fp, fp]
    L.mgpu_mlp_predict.argtypes = [E, dp, ctypes.c_longlong, fp]
    L.mgpu_timed_steps.argtypes = [E, ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mgpu_kernel_slot_name.argtypes = [E, ctypes.c_int]
    L.mgpu_kernel_slot_name.restype = ctypes.c_char_p
    tp = ctypes.POINTER(MgpuBatchTuning)
    L.mgpu_batch_create.argtypes = [ctypes.POINTER(MgpuBatchConfig), ctypes.POINTER(MgpuInstance), ctypes.c_int,
                                    ctypes.POINTER(E)]
    L.mgpu_batch_destroy.argtypes = [E]
    L.mgpu_batch_sizes.argtypes = [E, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.POINTER(ctypes.c_int)]
    L.mgpu_batch_set_prims.argtypes = [E, ctypes.c_int, dp]
    L.mgpu_batch_upload_state.argtypes = [E, ctypes.c_int, dp, dp, dp]
    L.mgpu_batch_download_state.argtypes = [E, ctypes.c_int, dp, dp, dp]
    L.mgpu_batch_step.argtypes = [E, ctypes.c_int, sp]
    L.mgpu_batch_get_time.argtypes = [E, ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_int)]
    L.mgpu_batch_use_graph.argtypes = [E, ctypes.c_int]
    L.mgpu_batch_get_tuning.argtypes = [E, tp]
    L.mgpu_batch_set_tuning.argtypes = [E, tp]
    L.mgpu_batch_autotune.argtypes = [E, ctypes.c_int, tp, tp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mgpu_batch_timed_steps.argtypes = [E, ctypes.c_int, dp]
    L.mgpu_batch_sync.argtypes = [E]
    _lib = L
    return L


def _fp(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Config:
    """Grid and scheme of one simulation (the reference Data constructor,
    simData.h:205-218, plus the model / WENO / integrator / Bcs choice)."""

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
    integrator: int = RK2
    bc: tuple = (PERIODIC, OUTFLOW, PERIODIC)
    quirks: int = 1
    device: int = 0

    def to_c(self) -> MgpuConfig:
        c = MgpuConfig()
        for f in ("nx", "ny", "nz", "Ng", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "gamma", "sigma",
                  "cp", "cfl", "model", "weno_order", "integrator", "quirks", "device"):
            setattr(c, f, getattr(self, f))
        for a in range(3):
            c.bc[a] = int(self.bc[a])
        return c

    @property
    def shape(self):
        return (self.nx + 2 * self.Ng, self.ny + 2 * self.Ng if self.ny > 0 else 1,
                self.nz + 2 * self.Ng if self.nz > 0 else 1)

    @property
    def interior_cells(self) -> int:
        return self.nx * max(self.ny, 1) * max(self.nz, 1)


class Engine:
    """One device-resident simulation driven through the C ABI."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.L = load_library()
        self._h = ctypes.c_void_p()
        c = cfg.to_c()
        self._check(self.L.mgpu_create(ctypes.byref(c), ctypes.byref(self._h)), "mgpu_create")
        nc, np_, na = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.L.mgpu_sizes(self._h, ctypes.byref(nc), ctypes.byref(np_), ctypes.byref(na)), "mgpu_sizes")
        self.ncons, self.nprims, self.naux = nc.value, np_.value, na.value

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise EngineError(f"{what}: {self.L.mgpu_last_error().decode()}")
        return rc

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.L.mgpu_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state ---------------------------------------------------------------
    def zeros(self, nvar: int) -> np.ndarray:
        return np.zeros((nvar,) + self.cfg.shape)

    def upload(self, cons=None, prims=None, aux=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (cons, prims, aux)]
        self._check(self.L.mgpu_upload_state(self._h, *[_dp(a) for a in arrs]), "mgpu_upload_state")

    def download(self):
        cons, prims, aux = self.zeros(self.ncons), self.zeros(self.nprims), self.zeros(self.naux)
        self._check(self.L.mgpu_download_state(self._h, _dp(cons), _dp(prims), _dp(aux)), "mgpu_download_state")
        return cons, prims, aux

    def set_prims(self, prims: np.ndarray):
        """Simulation::set: upload primitives, primsToAll, boundaries (simulation.cc:94-96)."""
        self.upload(prims=prims)
        self._check(self.L.mgpu_prims_to_all(self._h), "mgpu_prims_to_all")

    # time stepping -------------------------------------------------------
    def step(self, dt: float, stats: bool = False):
        st = MgpuStats() if stats else None
        rc = self._check(self.L.mgpu_step(self._h, dt, ctypes.byref(st) if st else None), "mgpu_step")
        return (rc, st.as_dict()) if stats else rc

    def update_time(self, end_time: float = 1e10, stats: bool = False):
        st = MgpuStats() if stats else None
        rc = self._check(self.L.mgpu_update_time(self._h, end_time, ctypes.byref(st) if st else None),
                         "mgpu_update_time")
        return (rc, st.as_dict()) if stats else rc

    def time(self):
        t, dt, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.L.mgpu_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(it)), "get_time")
        return t.value, dt.value, it.value

    def set_time(self, t: float, iters: int):
        self._check(self.L.mgpu_set_time(self._h, t, iters), "mgpu_set_time")

    def sync(self):
        self._check(self.L.mgpu_sync(self._h), "mgpu_sync")

    def set_sweep_plan(self, plan):
        """3-D grids: 3 (default): an x sweep, then y and z in one launch that
        finishes the stage; 2: sweeps x, y, z (the contiguous z finishes), the
        same bits as 3; 0: sweeps z, y, x (x finishes); 1 (or True): fused z+y
        launch, then x, the same bits as 0. 2-D grids: 3 runs y and x as one
        launch that finishes the stage, 0-2 a y sweep and an x sweep, all with
        the same bits."""
        self._check(self.L.mgpu_set_sweep_plan(self._h, int(plan)), "mgpu_set_sweep_plan")

    def set_regime_plan(self, fused: bool):
        """REGIME source: D formed on chip in the differencing launch (default)
        or stored and differenced by a second launch; bit-identical."""
        self._check(self.L.mgpu_set_regime_plan(self._h, 1 if fused else 0), "mgpu_set_regime_plan")

    # model / flux-method calls on the resident state ----------------------
    def fnet(self) -> np.ndarray:
        out = self.zeros(self.ncons)
        self._check(self.L.mgpu_fnet(self._h, _dp(out)), "mgpu_fnet")
        return out

    def flux_vector(self, dir: int) -> np.ndarray:
        out = self.zeros(self.ncons)
        self._check(self.L.mgpu_flux_vector(self._h, dir, _dp(out)), "mgpu_flux_vector")
        return out

    def source(self) -> np.ndarray:
        out = self.zeros(self.ncons)
        self._check(self.L.mgpu_source(self._h, _dp(out)), "mgpu_source")
        return out

    def flux_reconstruction(self, dir: int, cons=None, f=None, vars: int = -1) -> np.ndarray:
        """FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56): face fluxes at
        i - 1/2 along `dir` of given padded (cons, f), or of the resident state
        and its flux vector when both are None."""
        nv = self.ncons if vars < 0 else vars
        out = self.zeros(self.ncons)
        c = None if cons is None else np.ascontiguousarray(cons, dtype=np.float64)
        ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        self._check(self.L.mgpu_flux_reconstruction(self._h, _dp(c), _dp(ff), dir, vars, _dp(out)),
                    "mgpu_flux_reconstruction")
        return out[:nv]

    def c2p_cells(self, cons: np.ndarray, prims: np.ndarray, aux: np.ndarray):
        """Model::getPrimitiveVarsSingleCell over n cells ([cell][var] arrays);
        prims and aux are the starting guess. Returns (rc, prims, aux, stats)."""
        c = np.ascontiguousarray(cons, dtype=np.float64)
        p = np.array(prims, dtype=np.float64, order="C", copy=True)
        a = np.array(aux, dtype=np.float64, order="C", copy=True)
        n = c.shape[0]
        assert c.shape == (n, self.ncons) and p.shape == (n, self.nprims) and a.shape == (n, self.naux)
        st = MgpuStats()
        rc = self._check(self.L.mgpu_c2p_cells(self._h, n, _dp(c), _dp(p), _dp(a), ctypes.byref(st)),
                         "mgpu_c2p_cells")
        return rc, p, a, st.as_dict()

    def c2p_cells_noalloc(self, n: int, cons_ptr: int, prims_ptr: int, aux_ptr: int, stats: bool = True):
        """mgpu_c2p_cells on DEVICE arrays (addresses of [cell][var] arrays on
        this engine's GPU), allocating nothing. Returns (rc, stats or None);
        without stats the call only launches (sync before reading)."""
        st = MgpuStats() if stats else None
        rc = self._check(self.L.mgpu_c2p_cells_noalloc(self._h, int(n), ctypes.c_void_p(cons_ptr),
                                                       ctypes.c_void_p(prims_ptr), ctypes.c_void_p(aux_ptr),
                                                       ctypes.byref(st) if stats else None),
                         "mgpu_c2p_cells_noalloc")
        return rc, (st.as_dict() if stats else None)

    def source_cells(self, cons: np.ndarray, aux=None) -> np.ndarray:
        """Model::sourceTermSingleCell over n cells ([cell][var] arrays)."""
        c = np.ascontiguousarray(cons, dtype=np.float64)
        a = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
        out = np.zeros_like(c)
        self._check(self.L.mgpu_source_cells(self._h, c.shape[0], _dp(c), None, _dp(a), _dp(out)),
                    "mgpu_source_cells")
        return out

    def set_model_extension(self, kind: int):
        """Hand a model extension (EXT_REGIME) to the integrator, as the
        reference's REGIME(data, fluxMethod) passed to RKSplit / RKPlus."""
        self._check(self.L.mgpu_set_model_extension(self._h, kind), "mgpu_set_model_extension")

    def source_extension(self) -> np.ndarray:
        """REGIME::sourceExtension of the resident state (interior cells)."""
        out = self.zeros(self.ncons)
        self._check(self.L.mgpu_source_extension(self._h, _dp(out)), "mgpu_source_extension")
        return out

    def regime_fields(self):
        """REGIME's per-cell fields over the padded grid: E (3), q, alpha,
        dwdsb (12 x 3) and K (3), each with the grid shape after its
        components."""
        shape = self.cfg.shape
        E, K = np.zeros((3,) + shape), np.zeros((3,) + shape)
        q, alpha = np.zeros(shape), np.zeros(shape)
        dw = np.zeros((12, 3) + shape)
        self._check(self.L.mgpu_regime_fields(self._h, _dp(E), _dp(q), _dp(alpha), _dp(dw), _dp(K)),
                    "mgpu_regime_fields")
        return E, q, alpha, dw, K

    def c2p(self):
        st = MgpuStats()
        rc = self._check(self.L.mgpu_c2p(self._h, ctypes.byref(st)), "mgpu_c2p")
        return rc, st.as_dict()

    def max_wavespeed(self) -> np.ndarray:
        """Per-axis largest |characteristic speed| over the interior cells."""
        out = np.zeros(3)
        self._check(self.L.mgpu_max_wavespeed(self._h, _dp(out)), "mgpu_max_wavespeed")
        return out

    def timed_c2p(self, reps: int = 5):
        """(ms per cons2prim kernel call, counters of the last call), each call
        starting from the prims resident on entry."""
        ms = ctypes.c_double()
        st = MgpuStats()
        rc = self._check(self.L.mgpu_timed_c2p(self._h, reps, ctypes.byref(ms), ctypes.byref(st)), "mgpu_timed_c2p")
        return ms.value, st.as_dict(), rc

    def set_c2p_histogram(self, on: bool):
        """Start (clearing the 64 bins) or stop counting Newton steps per cell
        in c2p() / timed_c2p()."""
        self._check(self.L.mgpu_set_c2p_histogram(self._h, 1 if on else 0), "mgpu_set_c2p_histogram")

    def c2p_histogram(self) -> np.ndarray:
        """Cells per Newton-step count (bin 63: 63 or more, or not converged)."""
        h = np.zeros(64, dtype=np.int64)
        self._check(self.L.mgpu_get_c2p_histogram(self._h, h.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))),
                    "mgpu_get_c2p_histogram")
        return h

    def set_c2p_guess(self, mode: int, params=None, ymean_ystd=None):
        """Starting guess of c2p(): GUESS_PLAIN (previous primitives, the
        reference's), GUESS_LEARNED (network parameters from Mlp.params()) or
        GUESS_COLD."""
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        n = None if ymean_ystd is None else np.ascontiguousarray(ymean_ystd, dtype=np.float32)
        self._check(self.L.mgpu_set_c2p_guess(self._h, mode, _fp(p), _fp(n)), "mgpu_set_c2p_guess")

    # timing ----------------------------------------------------------------
    def timed_steps(self, dts, max_slots: int = 32):
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        total = ctypes.c_double()
        kms = np.zeros(max_slots)
        ns = ctypes.c_int()
        self._check(self.L.mgpu_timed_steps(self._h, len(dts), _dp(dts), ctypes.byref(total), _dp(kms), max_slots,
                                            ctypes.byref(ns)), "mgpu_timed_steps")
        slots = {}
        for i in range(ns.value):
            name = self.L.mgpu_kernel_slot_name(self._h, i).decode()
            if name != "start":
                slots[name] = kms[i]
        return total.value, slots


class Mlp:
    """The learned cons2prim guess network, trained on the GPU (C ABI mgpu_mlp_*)."""

    def __init__(self, seed: int = 0, device: int = 0):
        self.L = load_library()
        self._h = ctypes.c_void_p()
        rc = self.L.mgpu_mlp_create(device, seed, ctypes.byref(self._h))
        if rc < 0:
            raise EngineError(f"mgpu_mlp_create: {self.L.mgpu_last_error().decode()}")

    def _check(self, rc, what):
        if rc < 0:
            raise EngineError(f"{what}: {self.L.mgpu_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.L.mgpu_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_data(self, cons8: np.ndarray, w_p: np.ndarray):
        cons8 = np.ascontiguousarray(cons8, dtype=np.float64)
        w_p = np.ascontiguousarray(w_p, dtype=np.float64)
        assert cons8.shape[1] == 8 and w_p.shape == (cons8.shape[0], 2)
        self._check(self.L.mgpu_mlp_set_data(self._h, _dp(cons8), _dp(w_p), cons8.shape[0]), "mgpu_mlp_set_data")

    def train(self, steps: int, batch: int, lr: float) -> float:
        loss = ctypes.c_float()
        self._check(self.L.mgpu_mlp_train(self._h, steps, batch, lr, ctypes.byref(loss)), "mgpu_mlp_train")
        return loss.value

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.L.mgpu_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0], _fp(g), _fp(loss)),
                    "mgpu_mlp_gradients")
        return g, float(loss[0])

    def params(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        n = np.zeros(4, dtype=np.float32)
        self._check(self.L.mgpu_mlp_get_params(self._h, _fp(p), _fp(n)), "mgpu_mlp_get_params")
        return p, n

    def set_params(self, p, n=None):
        p = np.ascontiguousarray(p, dtype=np.float32)
        n = None if n is None else np.ascontiguousarray(n, dtype=np.float32)
        self._check(self.L.mgpu_mlp_set_params(self._h, _fp(p), _fp(n)), "mgpu_mlp_set_params")

    def predict(self, cons8: np.ndarray) -> np.ndarray:
        cons8 = np.ascontiguousarray(cons8, dtype=np.float64)
        y = np.zeros((cons8.shape[0], 2), dtype=np.float32)
        self._check(self.L.mgpu_mlp_predict(self._h, _dp(cons8), cons8.shape[0], _fp(y)), "mgpu_mlp_predict")
        return y


@dataclass
class Instance:
    """One member of a batched ensemble: a 2-D grid, its Courant number and
    the time at which it stops stepping."""

    nx: int
    ny: int
    xmin: float = -0.5
    xmax: float = 0.5
    ymin: float = -1.0
    ymax: float = 1.0
    cfl: float = 0.4
    end_time: float = 1e10

    def to_c(self) -> MgpuInstance:
        c = MgpuInstance()
        for f in ("nx", "ny", "xmin", "xmax", "ymin", "ymax", "cfl", "end_time"):
            setattr(c, f, getattr(self, f))
        return c


class Batch:
    """Independent 2-D ideal SRMHD simulations advanced together (C ABI mgpu_batch_*)."""

    def __init__(self, instances, weno_order: int = 5, bc=(PERIODIC, OUTFLOW), gamma: float = 4.0 / 3.0,
                 Ng: int = 4, quirks: int = 1, device: int = 0):
        self.L = load_library()
        self.instances = list(instances)
        self.Ng = Ng
        c = MgpuBatchConfig()
        c.Ng, c.gamma, c.weno_order, c.quirks, c.device = Ng, gamma, weno_order, quirks, device
        c.bc[0], c.bc[1] = int(bc[0]), int(bc[1])
        arr = (MgpuInstance * len(self.instances))(*[i.to_c() for i in self.instances])
        self._h = ctypes.c_void_p()
        self._check(self.L.mgpu_batch_create(ctypes.byref(c), arr, len(self.instances), ctypes.byref(self._h)),
                    "mgpu_batch_create")

    def _check(self, rc, what):
        if rc < 0:
            raise EngineError(f"{what}: {self.L.mgpu_last_error().decode()}")
        return rc

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.L.mgpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sizes(self):
        n, c, t = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()
        self._check(self.L.mgpu_batch_sizes(self._h, ctypes.byref(n), ctypes.byref(c), ctypes.byref(t)), "sizes")
        return n.value, c.value, t.value

    def shape(self, i: int):
        inst = self.instances[i]
        return (inst.nx + 2 * self.Ng, inst.ny + 2 * self.Ng, 1)

    def set_prims(self, i: int, prims: np.ndarray):
        prims = np.ascontiguousarray(prims, dtype=np.float64)
        assert prims.shape == (8,) + self.shape(i), prims.shape
        self._check(self.L.mgpu_batch_set_prims(self._h, i, _dp(prims)), "mgpu_batch_set_prims")

    def upload(self, i: int, cons, prims, aux):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (cons, prims, aux)]
        self._check(self.L.mgpu_batch_upload_state(self._h, i, *[_dp(a) for a in arrs]), "mgpu_batch_upload_state")

    def download(self, i: int):
        sh = self.shape(i)
        cons, prims, aux = np.zeros((9,) + sh), np.zeros((8,) + sh), np.zeros((13,) + sh)
        self._check(self.L.mgpu_batch_download_state(self._h, i, _dp(cons), _dp(prims), _dp(aux)),
                    "mgpu_batch_download_state")
        return cons, prims, aux

    def step(self, nsteps: int = 1, stats: bool = False):
        st = MgpuStats() if stats else None
        rc = self._check(self.L.mgpu_batch_step(self._h, nsteps, ctypes.byref(st) if st else None), "mgpu_batch_step")
        return (rc, st.as_dict()) if stats else rc

    def time(self, i: int):
        t, dt, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.L.mgpu_batch_get_time(self._h, i, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(it)),
                    "mgpu_batch_get_time")
        return t.value, dt.value, it.value

    def use_graph(self, on: bool):
        self._check(self.L.mgpu_batch_use_graph(self._h, 1 if on else 0), "mgpu_batch_use_graph")

    def tuning(self) -> dict:
        t = MgpuBatchTuning()
        self._check(self.L.mgpu_batch_get_tuning(self._h, ctypes.byref(t)), "mgpu_batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, tile_x: int, tile_y: int, schedule: int, blocks: int):
        t = MgpuBatchTuning(tile_x, tile_y, schedule, blocks, 0.0)
        self._check(self.L.mgpu_batch_set_tuning(self._h, ctypes.byref(t)), "mgpu_batch_set_tuning")

    def autotune(self, trial_steps: int = 3):
        """(chosen tuning, every candidate tried) -- the state is left as it was."""
        best = MgpuBatchTuning()
        cands = (MgpuBatchTuning * 256)()
        n = ctypes.c_int()
        self._check(self.L.mgpu_batch_autotune(self._h, trial_steps, ctypes.byref(best), cands, 256, ctypes.byref(n)),
                    "mgpu_batch_autotune")
        return best.as_dict(), [cands[k].as_dict() for k in range(min(n.value, 256))]

    def timed_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        self._check(self.L.mgpu_batch_timed_steps(self._h, nsteps, ctypes.byref(ms)), "mgpu_batch_timed_steps")
        return ms.value

    def sync(self):
        self._check(self.L.mgpu_batch_sync(self._h), "mgpu_batch_sync")
