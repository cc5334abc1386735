"""Python handles on the device-resident run and the ensemble
(include/lpbf_run.h).

DeviceRun mirrors lpbf::DeviceRun: the field stays on the GPU, steps run in
batches with one host synchronisation per batch, and per-step records
(statistics, W, W_sub) plus the running maxima of the reference MetricsTask
(include/gpu_step_task.hpp:375-404) come back.  Ensemble mirrors
lpbf::Ensemble: many runs of one plate (the reference manifest sweep,
src/gpu/hh_gpu_main.cpp:234-290) stepped together, FP64 by one launch per
step for all, FP32 by the batched multi-instance kernel.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LpbfError, RunOptions, StepConst, StepRecord, lib


def _u8(a, n):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.uint8).reshape(n))
    return arr, arr.ctypes.data


class DeviceRun:
    def __init__(self, k: StepConst, *, Tm: float, Lx: float, Ly: float, Tamb: float,
                 plan: np.ndarray, x_csv=None, wrapped=None, dump=None, phys: int = 0,
                 precision: int = 64, bc: int = 0, T_bc: float = 300.0, batch: int = 64,
                 depth: int = 4, rows: int = 0, T_ref: float = -1.0,
                 keep_records: bool = True, graph: bool = True, f64_depth: int = 1,
                 cruise_cols: bool = False):
        plan = np.ascontiguousarray(plan, dtype=np.float64)
        n = int(plan.shape[0]) if plan.size else 0
        self._keep = [plan]
        xc = None
        if x_csv is not None:
            xc = np.ascontiguousarray(x_csv, dtype=np.float64).reshape(n)
            self._keep.append(xc)
        wr, wr_p = _u8(wrapped, n)
        du, du_p = _u8(dump, n)
        self._keep += [wr, du]
        opt = RunOptions(precision, bc, T_bc, batch, depth, rows, T_ref, int(keep_records),
                         int(graph), f64_depth, int(cruise_cols), 0)
        h = C.c_void_p()
        rc = lib().lpbf_run_create(k, phys, Tm, Lx, Ly, Tamb, plan.ctypes.data if n else None,
                                   xc.ctypes.data if xc is not None else None, wr_p, du_p, n,
                                   C.byref(opt), C.byref(h))
        lib().lpbf_get_last_error()
        if rc != 0:
            raise LpbfError(f"lpbf_run_create: {lib().lpbf_error_string(rc).decode()}")
        self._h = h
        self.nx, self.ny, self.n_steps = k.nx, k.ny, n

    def _ok(self, rc, what):
        if rc != 0:
            lib().lpbf_get_last_error()
            raise LpbfError(f"{what}: {lib().lpbf_error_string(rc).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().lpbf_run_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, T: np.ndarray):
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(self.ny, self.nx)
        self._ok(lib().lpbf_run_upload(self._h, T.ctypes.data), "lpbf_run_upload")

    def noise(self, seed: int, base: float, amp: float):
        self._ok(lib().lpbf_run_noise(self._h, seed, base, amp), "lpbf_run_noise")

    def advance(self, n: int) -> int:
        r = lib().lpbf_run_advance(self._h, n)
        if r < 0:
            self._ok(-r, "lpbf_run_advance")
        return r

    def download(self) -> np.ndarray:
        T = np.empty((self.ny, self.nx))
        self._ok(lib().lpbf_run_download(self._h, T.ctypes.data), "lpbf_run_download")
        return T

    def summary(self):
        """(peak_T, peak_W_um, steady_W_sub_um, steps_done)."""
        out = (C.c_double * 3)()
        done = C.c_int()
        self._ok(lib().lpbf_run_summary(self._h, out, C.byref(done)), "lpbf_run_summary")
        return out[0], out[1], out[2], done.value

    def graph_launches(self) -> int:
        return lib().lpbf_run_graph_launches(self._h)

    def records(self, first: int = 0, n: int | None = None):
        if n is None:
            n = self.summary()[3] - first
        buf = (StepRecord * max(n, 1))()
        self._ok(lib().lpbf_run_records(self._h, first, n, buf), "lpbf_run_records")
        return list(buf)[:n]


class Ensemble:
    """n runs of one plate, advanced together (lpbf_ensemble_*): FP64 (one
    ensemble launch per step) or, with precision=32, the batched FP32 kernel
    (up to `depth` steps per launch for all runs still running).

    runs: a list of dicts with keys k (StepConst), plan ((n_i, 3) laser
    ticks), and optionally x_csv, wrapped (cruise gate) and Tamb."""

    def __init__(self, runs, *, Tm: float, Lx: float, Ly: float, phys: int = 0, bc: int = 0,
                 T_bc: float = 300.0, batch: int = 64, rows: int = 0, keep_records: bool = True,
                 graph: bool = True, precision: int = 64, depth: int = 4, T_ref: float = -1.0,
                 cruise_cols: bool = False):
        n = len(runs)
        ks = (StepConst * n)(*[r["k"] for r in runs])
        lens = [int(np.asarray(r["plan"]).reshape(-1, 3).shape[0]) for r in runs]
        plan = np.ascontiguousarray(np.concatenate(
            [np.asarray(r["plan"], dtype=np.float64).reshape(-1, 3) for r in runs]))
        have_x = all(r.get("x_csv") is not None for r in runs)
        xc = np.ascontiguousarray(np.concatenate(
            [np.asarray(r["x_csv"], dtype=np.float64).reshape(m) for r, m in zip(runs, lens)])) \
            if have_x else None
        wr = np.ascontiguousarray(np.concatenate(
            [np.asarray(r.get("wrapped", np.zeros(m)), dtype=np.uint8).reshape(m)
             for r, m in zip(runs, lens)])) if have_x else None
        tamb = np.ascontiguousarray([float(r.get("Tamb", 300.0)) for r in runs])
        steps = np.ascontiguousarray(lens, dtype=np.int32)
        self._keep = [ks, plan, xc, wr, tamb, steps]
        opt = RunOptions(precision, bc, T_bc, batch, depth, rows, T_ref, int(keep_records),
                         int(graph), 0, int(cruise_cols), 0)
        h = C.c_void_p()
        rc = lib().lpbf_ensemble_create(
            ks, n, phys, Tm, Lx, Ly, tamb.ctypes.data, plan.ctypes.data if plan.size else None,
            xc.ctypes.data if xc is not None else None, wr.ctypes.data if wr is not None else None,
            steps.ctypes.data, C.byref(opt), C.byref(h))
        lib().lpbf_get_last_error()
        if rc != 0:
            raise LpbfError(f"lpbf_ensemble_create: {lib().lpbf_error_string(rc).decode()}")
        self._h = h
        self.n = n
        self.nx, self.ny = runs[0]["k"].nx, runs[0]["k"].ny
        self.n_steps = lens

    def _ok(self, rc, what):
        if rc != 0:
            lib().lpbf_get_last_error()
            raise LpbfError(f"{what}: {lib().lpbf_error_string(rc).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().lpbf_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, m: int, T: np.ndarray):
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(self.ny, self.nx)
        self._ok(lib().lpbf_ensemble_upload(self._h, m, T.ctypes.data), "lpbf_ensemble_upload")

    def advance(self, n: int) -> int:
        r = lib().lpbf_ensemble_advance(self._h, n)
        if r < 0:
            self._ok(-r, "lpbf_ensemble_advance")
        return r

    def download(self, m: int) -> np.ndarray:
        T = np.empty((self.ny, self.nx))
        self._ok(lib().lpbf_ensemble_download(self._h, m, T.ctypes.data),
                 "lpbf_ensemble_download")
        return T

    def summary(self, m: int):
        """Run m: (peak_T, peak_W_um, steady_W_sub_um, steps_done)."""
        out = (C.c_double * 3)()
        done = C.c_int()
        self._ok(lib().lpbf_ensemble_summary(self._h, m, out, C.byref(done)),
                 "lpbf_ensemble_summary")
        return out[0], out[1], out[2], done.value

    def graph_launches(self) -> int:
        return lib().lpbf_ensemble_graph_launches(self._h)

    def records(self, m: int, first: int = 0, n: int | None = None):
        if n is None:
            n = self.summary(m)[3] - first
        buf = (StepRecord * max(n, 1))()
        self._ok(lib().lpbf_ensemble_records(self._h, m, first, n, buf), "lpbf_ensemble_records")
        return list(buf)[:n]
