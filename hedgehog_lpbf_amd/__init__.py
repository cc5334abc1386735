"""MI355X-native z-averaged 2D FD heat step of hedgehog_lpbf.

The compute lives in ``lib/libhedgehog_lpbf_amd.so`` (hand-written HIP
kernels for gfx950 behind a C ABI, ``include/lpbf_kernels.h``).  This package
is a thin Python mirror of the host interface for tests and benchmarks:
configuration (reference ``params.hpp``), laser plans, and device-resident
drivers that call the C ABI on torch-allocated device memory and streams.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (LPBF_BC_DIRICHLET, LPBF_BC_MIRROR, LPBF_PHYS_BASELINE,  # noqa: F401
                   LPBF_PHYS_EXTENDED, LaserTick, LpbfError, Partial, RunOptions, StepConst,
                   StepRecord, StepStats, check, lib)
from .batch import F32Batch, InstanceSpec  # noqa: F401
from .run import DeviceRun, Ensemble  # noqa: F401

__all__ = [
    "Material", "Process", "Domain", "make_const", "dt_cfl", "dt_cfl_ext",
    "single_track_plan", "raster_plan", "centred_raster", "apply_case",
    "StepConst", "Partial", "LaserTick", "StepStats", "lib", "check", "LpbfError",
    "DeviceRun", "Ensemble", "RunOptions", "StepRecord", "F32Batch", "InstanceSpec",
    "LPBF_BC_MIRROR", "LPBF_BC_DIRICHLET", "LPBF_PHYS_BASELINE", "LPBF_PHYS_EXTENDED",
]


@dataclass
class Material:
    """Reference include/params.hpp:13-21."""
    k: float = 25.0
    rho: float = 8190.0
    cp: float = 565.0
    Tm: float = 1609.0
    eps: float = 0.40
    eta: float = 0.35

    def alpha(self) -> float:
        return self.k / (self.rho * self.cp)

    def arr(self) -> np.ndarray:
        return np.array([self.k, self.rho, self.cp, self.Tm, self.eps, self.eta])


@dataclass
class Process:
    """Reference include/params.hpp:23-33."""
    P: float = 285.0
    V: float = 0.96
    r0: float = 33.5e-6
    h: float = 40.0e-6
    hConv: float = 100.0
    dSub: float = 30.0e-6
    Tamb: float = 298.15
    Tsub: float = 298.15
    radiation: bool = True

    def arr(self) -> np.ndarray:
        return np.array([self.P, self.V, self.r0, self.h, self.hConv, self.dSub,
                         self.Tamb, self.Tsub])


@dataclass
class Domain:
    """Reference include/params.hpp:35-40."""
    Lx: float = 1.6e-3
    Ly: float = 0.6e-3
    nx: int = 240
    ny: int = 120

    def dx(self) -> float:
        return self.Lx / (self.nx - 1)

    def dy(self) -> float:
        return self.Ly / (self.ny - 1)


_CASES = {  # reference src/main.cpp:97-109
    "0": (285.0, 0.96, 33.5e-6), "1.1": (285.0, 0.96, 24.5e-6),
    "1.2": (285.0, 0.96, 41.0e-6), "2.1": (285.0, 1.20, 33.5e-6),
    "2.2": (285.0, 0.80, 33.5e-6), "3.1": (325.0, 0.96, 33.5e-6),
    "3.2": (245.0, 0.96, 33.5e-6),
}


def apply_case(case_id: str, pr: Process) -> Process:
    if case_id not in _CASES:
        raise ValueError(f"unknown case '{case_id}'")
    pr.P, pr.V, pr.r0 = _CASES[case_id]
    return pr


def _dp(a: np.ndarray):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def make_const(mat: Material, pr: Process, dom: Domain, dt: float) -> StepConst:
    k = StepConst()
    m, p = mat.arr(), pr.arr()
    lib().lpbf_make_const(_dp(m), _dp(p), int(pr.radiation), dom.Lx, dom.Ly, dom.nx, dom.ny,
                          dt, k)
    return k


def dt_cfl(mat: Material, dom: Domain) -> float:
    m = mat.arr()
    return lib().lpbf_dt_cfl(_dp(m), dom.Lx, dom.Ly, dom.nx, dom.ny)


def dt_cfl_ext(mat: Material, dom: Domain, kmult: float = 1.0) -> float:
    m = mat.arr()
    return lib().lpbf_dt_cfl_ext(_dp(m), dom.Lx, dom.Ly, dom.nx, dom.ny, kmult)


def single_track_plan(V: float, dt: float, Lx: float, Ly: float, n_steps: int):
    """(plan[n,3] = x_l, y_l, q_scale per step; x_csv[n]; wrapped[n])."""
    plan = np.zeros((n_steps, 3))
    x_csv = np.zeros(n_steps)
    wrapped = np.zeros(n_steps, dtype=np.int32)
    lib().lpbf_plan_single_track(V, dt, Lx, Ly, n_steps, plan.ctypes.data, x_csv.ctypes.data,
                                 wrapped.ctypes.data)
    return plan, x_csv, wrapped.astype(bool)


def raster_plan(x0, y0, length, hatch, n_tracks, V, dt, n_steps) -> np.ndarray:
    plan = np.zeros((n_steps, 3))
    lib().lpbf_plan_raster(x0, y0, length, hatch, n_tracks, V, dt, n_steps, plan.ctypes.data)
    return plan


def centred_raster(Lx, Ly, V, dt, n_steps, n_tracks, hatch):
    """Raster parameters filling n_steps of travel, centred in the plate."""
    length = (V * dt * n_steps - (n_tracks - 1) * hatch) / n_tracks
    return dict(x0=0.5 * (Lx - length), y0=0.5 * Ly - 0.5 * (n_tracks - 1) * hatch,
                length=length, hatch=hatch, n_tracks=n_tracks, V=V)
