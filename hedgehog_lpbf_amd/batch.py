"""Python handle on batched FP32 instances (include/lpbf_batch.h).

Torch allocates the device fields, plans and statistics rings; every step
runs in the library's multi-instance kernel.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import F32Instance, LpbfError, StepConst, StepStats, lib


@dataclass
class InstanceSpec:
    k: StepConst             # grid and physics
    plan: np.ndarray         # [n_steps, 3] laser (x_l, y_l, q_scale)
    T0: float | np.ndarray   # initial temperature (uniform or [ny, nx])
    T_ref: float = 300.0
    T_bc: float = 300.0
    Tm: float = 1609.0
    size_class: int = 0
    stats: bool = True


def _ok(rc, what):
    if rc != 0:
        lib().lpbf_get_last_error()
        raise LpbfError(f"{what}: {lib().lpbf_error_string(rc).decode()}")


class F32Batch:
    def __init__(self, specs: list[InstanceSpec], bc: int = 0, device="cuda"):
        import torch
        self._torch = torch
        self.specs = specs
        self.n = len(specs)
        self._bufs = []
        arr = (F32Instance * self.n)()
        for i, sp in enumerate(specs):
            nx, ny = sp.k.nx, sp.k.ny
            ld = (nx + 3) & ~3
            f0 = torch.zeros((ny, ld), dtype=torch.float32, device=device)
            f1 = torch.zeros_like(f0)
            T0 = np.broadcast_to(np.asarray(sp.T0, dtype=np.float64), (ny, nx))
            f0[:, :nx] = torch.from_numpy(np.ascontiguousarray(T0 - sp.T_ref)).to(
                device=device, dtype=torch.float32)
            plan = torch.from_numpy(np.ascontiguousarray(sp.plan, dtype=np.float64)).to(device)
            stats = (torch.zeros(plan.shape[0] * C.sizeof(StepStats), dtype=torch.uint8,
                                 device=device) if sp.stats else None)
            self._bufs.append((f0, f1, plan, stats))
            a = arr[i]
            a.k = StepConst.from_buffer_copy(bytes(sp.k))
            a.theta0, a.theta1, a.ld = f0.data_ptr(), f1.data_ptr(), ld
            a.size_class, a.T_ref, a.T_bc, a.Tm = sp.size_class, sp.T_ref, sp.T_bc, sp.Tm
            a.d_plan = plan.data_ptr()
            a.d_stats = stats.data_ptr() if stats is not None else None
            a.n_steps = plan.shape[0]
        torch.cuda.synchronize()
        h = C.c_void_p()
        _ok(lib().lpbf_f32_batch_create(arr, self.n, bc, C.byref(h)), "lpbf_f32_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lpbf_f32_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _stream(self, stream):
        return stream if stream is not None else self._torch.cuda.current_stream().cuda_stream

    def set_tiling(self, size_class: int, depth: int, rows: int = 0):
        _ok(lib().lpbf_f32_batch_set_tiling(self._h, size_class, depth, rows), "set_tiling")

    def tiling(self, size_class: int):
        d, r, m = C.c_int(), C.c_int(), C.c_int()
        _ok(lib().lpbf_f32_batch_get_tiling(self._h, size_class, C.byref(d), C.byref(r),
                                            C.byref(m)), "get_tiling")
        return d.value, r.value, m.value

    def autotune(self, max_depth: int = 4, stream=None) -> float:
        ms = C.c_double()
        _ok(lib().lpbf_f32_batch_autotune(self._h, max_depth, self._stream(stream), C.byref(ms)),
            "autotune")
        return ms.value

    def enqueue(self, step0: int, n_steps: int, stream=None):
        _ok(lib().lpbf_f32_batch_enqueue(self._h, step0, n_steps, self._stream(stream)), "enqueue")

    def capture(self, step0: int, n_steps: int, stream):
        _ok(lib().lpbf_f32_batch_capture(self._h, step0, n_steps, stream), "capture")

    def launch_graph(self, stream=None):
        _ok(lib().lpbf_f32_batch_launch_graph(self._h, self._stream(stream)), "launch_graph")

    def launches(self, n_steps: int):
        nl, cells = C.c_int(), C.c_longlong()
        _ok(lib().lpbf_f32_batch_launches(self._h, n_steps, C.byref(nl), C.byref(cells)),
            "launches")
        return nl.value, cells.value

    def theta(self, i: int):
        c = lib().lpbf_f32_batch_current(self._h, i)
        if c < 0:
            _ok(-c, "current")
        return self._bufs[i][c]

    def field(self, i: int) -> np.ndarray:
        """Latest T of instance i (FP64 numpy, [ny, nx])."""
        sp = self.specs[i]
        th = self.theta(i)[:, :sp.k.nx].double().cpu().numpy()
        return th + sp.T_ref

    def stats(self, i: int, first: int = 0, n: int | None = None):
        buf = self._bufs[i][3]
        if buf is None:
            return []
        total = buf.numel() // C.sizeof(StepStats)
        n = total - first if n is None else n
        raw = buf.cpu().numpy().tobytes()
        return list((StepStats * total).from_buffer_copy(raw))[first:first + n]
