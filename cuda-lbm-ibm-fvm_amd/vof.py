"""Python host interface of the MI355X VOF timestep (ctypes over include/vof_capi.h).

Mirrors the reference solver's surface (voftrial1.cu): a parameter block with
the reference defaults (``Params``/``cpar_host``, :27-41, :1258-1291), the
dam-break initial condition (:72-93), the time-step loop body (:1361-1401),
the download of F for output (:1405) and the legacy VTK writer (:907-941).

Everything runs in ``lib/libvof_hip.so`` (hand-written gfx950 HIP kernels).
There is no CPU fallback: if the library is missing, :func:`load_library`
raises.  Arrays cross the boundary in the reference layout: shape
(nz+2, ny+2, nx+2), x fastest, one ghost layer.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint32, c_void_p

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libvof_hip.so")
CLI_PATH = os.path.join(PKG, "lib", "3dvof")

FIELDS = {"F": 0, "U": 1, "V": 2, "W": 3, "P": 4, "USTAR": 5, "VSTAR": 6, "WSTAR": 7, "RHS": 8, "RHO": 9,
          "NU": 10, "KAPPA": 11}
FLAG_REMOVE_P_MEAN = 0x1
FLAG_CSF = 0x2
FLAG_RHO_PREDICTOR = 0x4
FLAG_NO_SLIP = 0x8
FLAG_WATCHDOG = 0x10

# Order of the per-stage times returned by Solver.time_steps(stages=True).
STAGE_TIME_NAMES = ("predictor", "rhs", "jacobi", "p_mean", "correction", "fct_x", "fct_y", "fct_z", "ghost_F",
                    "fct_zx", "fct_xy")

# Stages for Solver.debug_stage (test hook).
(STAGE_PREDICT, STAGE_RHS, STAGE_JACOBI, STAGE_CORRECT, STAGE_FCT_X, STAGE_FCT_Y, STAGE_FCT_Z, STAGE_GHOST_F,
 STAGE_JACOBI2, STAGE_CSF, STAGE_JACOBI_T) = range(11)


class VofParams(ctypes.Structure):
    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32), ("precision", c_int32),
        ("Lx", c_double), ("Ly", c_double), ("Lz", c_double),
        ("rho_l", c_double), ("rho_g", c_double), ("nu_l", c_double), ("nu_g", c_double),
        ("sigma", c_double), ("gx", c_double), ("gy", c_double), ("gz", c_double), ("dt", c_double),
        ("jacobi_sweeps", c_int32), ("flags", c_uint32), ("jacobi_tblock", c_int32), ("reserved", c_int32),
    ]


class VofDiag(ctypes.Structure):
    _fields_ = [
        ("liquid_volume", c_double), ("p_mean", c_double), ("residual_l2", c_double), ("rhs_l2", c_double),
        ("f_min", c_double), ("f_max", c_double), ("nonfinite", c_int64),
    ]


class VofAdjScalars(ctypes.Structure):
    _fields_ = [("d_rho_l", c_double), ("d_rho_g", c_double), ("d_nu_l", c_double), ("d_nu_g", c_double),
                ("d_gx", c_double), ("d_gy", c_double), ("d_gz", c_double), ("d_sigma", c_double)]


class VofTuning(ctypes.Structure):
    _fields_ = [("kz_predict", c_int32), ("kz_rhs", c_int32), ("kz_jacobi", c_int32), ("kz_correct", c_int32),
                ("fct_chunk", c_int32), ("jacobi_tblock", c_int32), ("kz_fct_x", c_int32),
                ("fct_zx_fused", c_int32), ("fct_xy_fused", c_int32)]


def default_params(**kw) -> VofParams:
    """Reference defaults (voftrial1.cu:1258-1283) with keyword overrides.

    ``n=`` sets nx = ny = nz; ``g=(gx, gy, gz)`` sets gravity.
    """
    p = VofParams()
    p.nx = p.ny = p.nz = 200
    p.precision = 4
    p.Lx = p.Ly = p.Lz = 0.1
    p.rho_l, p.rho_g = 1000.0, 50.0
    p.nu_l, p.nu_g = 1.0e-6, 1.5e-5
    p.sigma = 0.007
    p.gx, p.gy, p.gz = 0.0, -5.0, 0.0
    p.dt = 4.0e-6
    p.jacobi_sweeps = 10
    p.flags = 0
    p.jacobi_tblock = 0
    n = kw.pop("n", None)
    if n is not None:
        p.nx = p.ny = p.nz = int(n)
    g = kw.pop("g", None)
    if g is not None:
        p.gx, p.gy, p.gz = (float(x) for x in g)
    L = kw.pop("L", None)
    if L is not None:
        p.Lx = p.Ly = p.Lz = float(L)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k!r}")
        setattr(p, k, v)
    return p


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libvof_hip.so and declare its C ABI.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (there is no CPU fallback)")
    L = ctypes.CDLL(path)
    S, B = c_void_p, c_void_p
    sig = {
        "vof_params_default": (None, [POINTER(VofParams)]),
        "vof_create": (c_int, [POINTER(VofParams), POINTER(S)]),
        "vof_destroy": (None, [S]),
        "vof_init_ic": (c_int, [S, c_int]),
        "vof_init_box": (c_int, [S, c_double, c_double, c_double]),
        "vof_step": (c_int, [S, c_int64, c_int]),
        "vof_step_graph": (c_int, [S, c_int64, c_int]),
        "vof_download": (c_int, [S, c_int, c_void_p, c_size_t]),
        "vof_upload": (c_int, [S, c_int, c_void_p, c_size_t]),
        "vof_diagnostics": (c_int, [S, POINTER(VofDiag)]),
        "vof_diagnostics_noalloc": (c_int, [POINTER(VofParams), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p]),
        "vof_write_vtk": (c_int, [S, c_char_p]),
        "vof_write_vtk_async": (c_int, [S, c_char_p]),
        "vof_output_wait": (c_int, [S]),
        "vof_save_state": (c_int, [S, c_char_p, c_int64]),
        "vof_load_state": (c_int, [S, c_char_p, POINTER(c_int64)]),
        "vof_write_vtk_host": (c_int, [c_char_p, c_void_p, c_int, c_int, c_int, c_int]),
        "vof_debug_vtk_thread_limit": (None, [c_int]),
        "vof_copy_bandwidth": (c_int, [c_int64, c_int, POINTER(c_double)]),
        "vof_stream_bandwidth": (c_int, [c_int64, c_int, c_int, c_int, c_int, POINTER(c_double)]),
        "vof_sync": (c_int, [S]),
        "vof_time_steps": (c_int, [S, c_int64, c_int, c_int, c_int, POINTER(c_double), POINTER(c_double)]),
        "vof_set_tuning": (c_int, [S, POINTER(VofTuning)]),
        "vof_get_tuning": (c_int, [S, POINTER(VofTuning)]),
        "vof_debug_stage": (c_int, [S, c_int]),
        "vof_autotune": (c_int, [S, c_int64, c_int, c_char_p, POINTER(VofTuning)]),
        "vof_batch_autotune": (c_int, [B, c_int64, c_int, c_char_p, POINTER(VofTuning)]),
        "vof_interior_cells": (c_int64, [S]),
        "vof_total_cells": (c_int64, [S]),
        "vof_status_str": (c_char_p, [c_int]),
        "vof_last_error": (c_char_p, []),
        "vof_batch_create": (c_int, [POINTER(VofParams), c_int, POINTER(B)]),
        "vof_batch_destroy": (None, [B]),
        "vof_batch_init_box": (c_int, [B, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
        "vof_batch_init_ic": (c_int, [B, c_int]),
        "vof_batch_step": (c_int, [B, c_int64, c_int, c_int]),
        "vof_batch_download": (c_int, [B, c_int, c_int, c_void_p, c_size_t]),
        "vof_batch_upload": (c_int, [B, c_int, c_int, c_void_p, c_size_t]),
        "vof_batch_diagnostics": (c_int, [B, c_int, POINTER(VofDiag)]),
        "vof_batch_sync": (c_int, [B]),
        "vof_batch_time_steps": (c_int, [B, c_int64, c_int, c_int, c_int, POINTER(c_double)]),
        "vof_batch_set_tuning": (c_int, [B, POINTER(VofTuning)]),
        "vof_batch_interior_cells": (c_int64, [B]),
        "vof_batch_size": (c_int, [B]),
        "vof_adj_forward": (c_int, [S, c_int64, c_int, c_int, POINTER(c_double)]),
        "vof_adj_seed": (c_int, [S, c_int, c_void_p, c_size_t]),
        "vof_adj_backward": (c_int, [S, c_double, POINTER(VofAdjScalars)]),
        "vof_adj_download": (c_int, [S, c_int, c_void_p, c_size_t]),
        "vof_adj_times": (c_int, [S, POINTER(c_double), POINTER(c_double)]),
        "vof_adj_release": (c_int, [S]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class VofError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def _check(st: int) -> None:
    if st != 0:
        L = load_library()
        raise VofError(st, L.vof_last_error().decode())


def _dtype(prec: int):
    return np.float64 if prec == 8 else np.float32


class Solver:
    """One dam-break instance on the GPU (the reference's DeviceFields + step)."""

    def __init__(self, params: VofParams | None = None, **kw):
        self.lib = load_library()
        self.params = params if params is not None else default_params(**kw)
        h = c_void_p()
        _check(self.lib.vof_create(byref(self.params), byref(h)))
        self.h = h
        p = self.params
        self.shape = (p.nz + 2, p.ny + 2, p.nx + 2)
        self.dtype = _dtype(p.precision)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.vof_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_ic(self, ic: int = 1) -> None:
        _check(self.lib.vof_init_ic(self.h, ic))

    def init_box(self, bx: float, by: float, bz: float) -> None:
        _check(self.lib.vof_init_box(self.h, bx, by, bz))

    def step(self, first_istep: int, nsteps: int = 1, graph: bool = False) -> None:
        fn = self.lib.vof_step_graph if graph else self.lib.vof_step
        _check(fn(self.h, first_istep, nsteps))

    def debug_stage(self, stage: int) -> None:
        _check(self.lib.vof_debug_stage(self.h, stage))

    def download(self, field: str) -> np.ndarray:
        a = np.empty(self.shape, dtype=self.dtype)
        _check(self.lib.vof_download(self.h, FIELDS[field], a.ctypes.data, a.size))
        return a

    def upload(self, field: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _check(self.lib.vof_upload(self.h, FIELDS[field], a.ctypes.data, a.size))

    def diagnostics(self) -> dict:
        d = VofDiag()
        _check(self.lib.vof_diagnostics(self.h, byref(d)))
        return {k: getattr(d, k) for k, _ in VofDiag._fields_}

    def write_vtk(self, path: str) -> None:
        _check(self.lib.vof_write_vtk(self.h, path.encode()))

    def write_vtk_async(self, path: str) -> None:
        """The same file, written while later steps run (see output_wait)."""
        _check(self.lib.vof_write_vtk_async(self.h, path.encode()))

    def save_state(self, path: str, istep: int) -> None:
        """Restart file of the state after step `istep` (both F buffers, u, v, w, p)."""
        _check(self.lib.vof_save_state(self.h, path.encode(), int(istep)))

    def load_state(self, path: str) -> int:
        """Load a restart file; returns the step it was saved after."""
        ist = c_int64()
        _check(self.lib.vof_load_state(self.h, path.encode(), byref(ist)))
        return ist.value

    def output_wait(self) -> None:
        """Wait for every pending write_vtk_async file; raises its error if one failed."""
        _check(self.lib.vof_output_wait(self.h))

    def sync(self) -> None:
        _check(self.lib.vof_sync(self.h))

    def time_steps(self, first_istep: int, warmup: int, steps: int, graph: bool = False, stages: bool = False):
        """Device milliseconds for `steps` steps (HIP events on the solver's
        stream); with stages=True also a dict of per-stage milliseconds."""
        ms = c_double()
        st = (c_double * len(STAGE_TIME_NAMES))()
        _check(self.lib.vof_time_steps(self.h, first_istep, warmup, steps, int(graph), byref(ms),
                                       st if stages else None))
        return (ms.value, dict(zip(STAGE_TIME_NAMES, st))) if stages else ms.value

    def set_tuning(self, **kw) -> None:
        t = VofTuning(*(int(kw.get(k, 0)) for k, _ in VofTuning._fields_))
        _check(self.lib.vof_set_tuning(self.h, byref(t)))

    def get_tuning(self) -> dict:
        t = VofTuning()
        _check(self.lib.vof_get_tuning(self.h, byref(t)))
        return {k: getattr(t, k) for k, _ in VofTuning._fields_}

    def autotune(self, first_istep: int = 1, steps_per_trial: int = 5, cache: str | None = None) -> dict:
        """Pick launch shapes by timing candidates on the live state (restored
        afterwards); cached per (arch, grid, precision) in the JSON file `cache`."""
        t = VofTuning()
        _check(self.lib.vof_autotune(self.h, first_istep, steps_per_trial, cache.encode() if cache else None,
                                     byref(t)))
        return {k: getattr(t, k) for k, _ in VofTuning._fields_}

    @property
    def interior_cells(self) -> int:
        return int(self.lib.vof_interior_cells(self.h))

    # ---- reverse mode: objective J = liquid-centroid height after the last step
    def adj_forward(self, first_istep: int, nsteps: int, checkpoint_every: int = 20) -> float:
        """Advance nsteps from the current state keeping checkpoints; returns J."""
        J = c_double()
        _check(self.lib.vof_adj_forward(self.h, first_istep, nsteps, checkpoint_every, byref(J)))
        return J.value

    def adj_seed(self, field: str, arr: np.ndarray) -> None:
        """Extra FP64 seed of a final-state field (cleared by adj_forward)."""
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _check(self.lib.vof_adj_seed(self.h, FIELDS[field], a.ctypes.data, a.size))

    def adj_backward(self, gJ: float = 1.0) -> dict:
        """Reverse sweep; returns the parameter gradients (times gJ)."""
        s = VofAdjScalars()
        _check(self.lib.vof_adj_backward(self.h, float(gJ), byref(s)))
        return {k[2:]: getattr(s, k) for k, _ in VofAdjScalars._fields_}

    def adj_download(self, field: str) -> np.ndarray:
        """Adjoint (FP64) of an initial-state field: F, U, V, W or P."""
        a = np.empty(self.shape, dtype=np.float64)
        _check(self.lib.vof_adj_download(self.h, FIELDS[field], a.ctypes.data, a.size))
        return a

    def adj_times(self) -> tuple[float, float]:
        """Device milliseconds of the last adj_forward and adj_backward."""
        f, b = c_double(), c_double()
        _check(self.lib.vof_adj_times(self.h, byref(f), byref(b)))
        return f.value, b.value

    def adj_release(self) -> None:
        _check(self.lib.vof_adj_release(self.h))


class Batch:
    """Ensemble of instances of uneven size advanced with one launch per stage."""

    def __init__(self, params_list):
        self.lib = load_library()
        n = len(params_list)
        arr = (VofParams * n)(*params_list)
        h = c_void_p()
        _check(self.lib.vof_batch_create(arr, n, byref(h)))
        self.h = h
        self.params = list(params_list)
        self.n = n
        self.dtype = _dtype(params_list[0].precision)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.vof_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i: int):
        p = self.params[i]
        return (p.nz + 2, p.ny + 2, p.nx + 2)

    def init_box(self, bx, by, bz) -> None:
        mk = lambda v: (c_double * self.n)(*[float(x) for x in v])
        _check(self.lib.vof_batch_init_box(self.h, mk(bx), mk(by), mk(bz)))

    def init_ic(self, ic: int = 1) -> None:
        _check(self.lib.vof_batch_init_ic(self.h, ic))

    def step(self, first_istep: int, nsteps: int = 1, graph: bool = False) -> None:
        _check(self.lib.vof_batch_step(self.h, first_istep, nsteps, int(graph)))

    def download(self, i: int, field: str) -> np.ndarray:
        a = np.empty(self.shape(i), dtype=self.dtype)
        _check(self.lib.vof_batch_download(self.h, i, FIELDS[field], a.ctypes.data, a.size))
        return a

    def upload(self, i: int, field: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        _check(self.lib.vof_batch_upload(self.h, i, FIELDS[field], a.ctypes.data, a.size))

    def diagnostics(self, i: int) -> dict:
        d = VofDiag()
        _check(self.lib.vof_batch_diagnostics(self.h, i, byref(d)))
        return {k: getattr(d, k) for k, _ in VofDiag._fields_}

    def sync(self) -> None:
        _check(self.lib.vof_batch_sync(self.h))

    def time_steps(self, first_istep: int, warmup: int, steps: int, graph: bool = False) -> float:
        ms = c_double()
        _check(self.lib.vof_batch_time_steps(self.h, first_istep, warmup, steps, int(graph), byref(ms)))
        return ms.value

    def set_tuning(self, **kw) -> None:
        t = VofTuning(*(int(kw.get(k, 0)) for k, _ in VofTuning._fields_))
        _check(self.lib.vof_batch_set_tuning(self.h, byref(t)))

    def autotune(self, first_istep: int = 1, steps_per_trial: int = 3, cache: str | None = None) -> dict:
        t = VofTuning()
        _check(self.lib.vof_batch_autotune(self.h, first_istep, steps_per_trial,
                                           cache.encode() if cache else None, byref(t)))
        return {k: getattr(t, k) for k, _ in VofTuning._fields_}

    @property
    def interior_cells(self) -> int:
        return int(self.lib.vof_batch_interior_cells(self.h))


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Attainable device bandwidth in GB/s (read + write) of a plain copy kernel."""
    lib = load_library()
    g = c_double()
    _check(lib.vof_copy_bandwidth(int(nbytes), int(reps), byref(g)))
    return g.value


def stream_time(field_bytes: int, nread: int, nwrite: int, resident: bool = False, reps: int = 20) -> float:
    """Microseconds one pass of a plain stream kernel needs to read `nread` and
    write `nwrite` arrays of `field_bytes` bytes each (inputs from HBM unless
    `resident`): the attainable time of a step kernel's compulsory traffic."""
    lib = load_library()
    us = c_double()
    _check(lib.vof_stream_bandwidth(int(field_bytes), int(nread), int(nwrite), int(bool(resident)), int(reps),
                                    byref(us)))
    return us.value


def write_vtk_host(path: str, field: np.ndarray) -> None:
    """Legacy VTK of a host array (same bytes as Solver.write_vtk)."""
    lib = load_library()
    a = np.ascontiguousarray(field)
    prec = 8 if a.dtype == np.float64 else 4
    a = a.astype(np.float64 if prec == 8 else np.float32, copy=False)
    nz, ny, nx = a.shape
    _check(lib.vof_write_vtk_host(path.encode(), a.ctypes.data, prec, nx, ny, nz))
