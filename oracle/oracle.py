"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / reported baseline, never
as the product path.  The oracle restates voftrial1.cu's kernel bodies as
OpenMP host loops (see vof_oracle.cpp for the line-by-line citations).
"""
from __future__ import annotations

import ctypes
import os
import sys
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_longlong, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# VOF_ORACLE_LIB names another build of the same source: the mutation tests
# (tests/test_oracle_mutants.py) point it at copies with a planted mistake.
LIB_PATH = os.environ.get("VOF_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
sys.path.insert(0, os.path.join(ROOT, "cuda-lbm-ibm-fvm_amd"))
from vof import VofDiag, VofParams, default_params  # noqa: E402  (struct layouts only)

# Oracle array ids for the scratch arrays (raw access: 100 + id).
RAW = {"F": 0, "FTD": 1, "AX": 2, "AY": 3, "AZ": 4, "CX": 5, "CY": 6, "CZ": 7, "RP": 8, "RM": 9,
       "U": 10, "V": 11, "W": 12, "USTAR": 13, "VSTAR": 14, "WSTAR": 15, "P": 16, "PT": 17, "RHO": 18, "NU": 19,
       "FSM": 20, "KAPPA": 21, "S1": 22, "S4": 23}
FIELDS = {"F": 0, "U": 1, "V": 2, "W": 3, "P": 4, "USTAR": 5, "VSTAR": 6, "WSTAR": 7, "RHS": 8, "RHO": 9,
          "NU": 10, "KAPPA": 11}
STAGES = {"nu_rho": 0, "advect_u": 1, "advect_v": 2, "advect_w": 3, "jacobi": 4, "update": 5, "bc_vel": 6,
          "fct_x": 7, "fct_y": 8, "fct_z": 9, "clamp": 10, "bc_F": 11, "remove_p_mean": 12,
          "curvature": 13, "csf": 14}

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing; build it with cuda-lbm-ibm-fvm_amd/build_native.py")
    L = ctypes.CDLL(LIB_PATH)
    L.orc_create.restype = c_void_p
    L.orc_create.argtypes = [POINTER(VofParams)]
    L.orc_destroy.argtypes = [c_void_p]
    L.orc_init_ic.argtypes = [c_void_p, c_int]
    L.orc_init_box.argtypes = [c_void_p, c_double, c_double, c_double]
    L.orc_step.argtypes = [c_void_p, c_longlong, c_int]
    L.orc_stage.argtypes = [c_void_p, c_int]
    L.orc_get.argtypes = [c_void_p, c_int, c_void_p]
    L.orc_set.argtypes = [c_void_p, c_int, c_void_p]
    L.orc_get_raw.argtypes = [c_void_p, c_int, c_void_p]
    L.orc_diag.argtypes = [c_void_p, POINTER(VofDiag)]
    L.orc_total.restype = c_longlong
    L.orc_total.argtypes = [c_void_p]
    L.orc_clamp01f.restype = c_float
    L.orc_clamp01f.argtypes = [c_float]
    L.orc_clamp01d.restype = c_double
    L.orc_clamp01d.argtypes = [c_double]
    L.orc_write_vtk_float.restype = c_int
    L.orc_write_vtk_float.argtypes = [c_char_p, c_void_p, c_int, c_int, c_int]
    L.orc_threads.restype = c_int
    L.orc_stats_enable.argtypes = [c_void_p, c_int]
    L.orc_stats.argtypes = [c_void_p, c_void_p]
    _lib = L
    return L


def threads() -> int:
    """OpenMP threads used by the oracle's loops."""
    return int(load().orc_threads())


def clamp01f(v: float) -> float:
    return float(load().orc_clamp01f(v))


def clamp01d(v: float) -> float:
    return float(load().orc_clamp01d(v))


def write_vtk_float(path: str, field: np.ndarray) -> None:
    a = np.ascontiguousarray(field, dtype=np.float32)
    nz, ny, nx = a.shape
    if load().orc_write_vtk_float(path.encode(), a.ctypes.data, nx, ny, nz) != 0:
        raise IOError(path)


class Oracle:
    """CPU restatement of the reference step (float or double)."""

    def __init__(self, params: VofParams | None = None, **kw):
        self.lib = load()
        self.params = params if params is not None else default_params(**kw)
        self.h = self.lib.orc_create(byref(self.params))
        if not self.h:
            raise RuntimeError("orc_create failed")
        p = self.params
        self.shape = (p.nz + 2, p.ny + 2, p.nx + 2)
        self.dtype = np.float64 if p.precision == 8 else np.float32

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_ic(self, ic: int = 1):
        self.lib.orc_init_ic(self.h, ic)

    def init_box(self, bx, by, bz):
        self.lib.orc_init_box(self.h, bx, by, bz)

    def step(self, first_istep: int, nsteps: int = 1):
        self.lib.orc_step(self.h, first_istep, nsteps)

    def stage(self, name: str):
        if self.lib.orc_stage(self.h, STAGES[name]) != 0:
            raise ValueError(name)

    def get(self, field: str) -> np.ndarray:
        """Field in its native precision (bit-exact copy)."""
        a = np.empty(self.shape, dtype=self.dtype)
        if field == "RHS":
            d = np.empty(self.shape, dtype=np.float64)
            self.lib.orc_get(self.h, FIELDS["RHS"], d.ctypes.data)
            return d.astype(self.dtype)
        idx = FIELDS[field] if field in FIELDS else 100 + RAW[field]
        if field in ("FTD", "AX", "AY", "AZ", "CX", "CY", "CZ", "RP", "RM", "PT"):
            idx = 100 + RAW[field]
        if self.lib.orc_get_raw(self.h, idx, a.ctypes.data) != 0:
            raise ValueError(field)
        return a

    def set(self, field: str, arr: np.ndarray):
        d = np.ascontiguousarray(arr, dtype=np.float64)
        idx = FIELDS.get(field, None)
        if idx is None or field in ("RHS",):
            idx = 100 + RAW[field]
        if self.lib.orc_set(self.h, idx, d.ctypes.data) != 0:
            raise ValueError(field)

    def stats_enable(self, on: bool = True):
        """Start (and reset) the liquid-volume accounting of later steps."""
        self.lib.orc_stats_enable(self.h, int(on))

    def stats(self) -> dict:
        """Volume changes summed since stats_enable, in m^3: clip1 (stage-1
        clip), clamp4 (stage-4 clamp01: clipping + rounding), clip4 (its plain
        clipping part), dil (sum F (vol - dv) of the sweeps), post
        (post_process_f's clamp01), div (dt vol sum F div u per step)."""
        out = np.zeros(6)
        self.lib.orc_stats(self.h, out.ctypes.data)
        return dict(zip(("clip1", "clamp4", "clip4", "dil", "post", "div"), out.tolist()))

    def diagnostics(self) -> dict:
        d = VofDiag()
        self.lib.orc_diag(self.h, byref(d))
        return {k: getattr(d, k) for k, _ in VofDiag._fields_}
