"""Melt-pool surrogate: training data from the FD solver, training with the
library's BF16-MFMA forward/backward kernels and Adam, and the
surrogate-accelerated run with FD recheck and fallback
(include/lpbf_surrogate.h).  Torch holds the device tensors; every
convolution, gradient and optimizer step runs in libhedgehog_lpbf_amd.so.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from ._lib import LPBF_BC_MIRROR, LaserTick, LpbfError, StepConst, check, lib

WIN = 128
CH = 32
NPARAM = 38465
STEPS = 16


class SgWindow(C.Structure):
    """include/lpbf_surrogate.h lpbf_sg_window."""
    _fields_ = [("theta", C.c_void_p), ("theta_out", C.c_void_p), ("plan", C.c_void_p),
                ("T_ref", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("inv_r0_sq", C.c_double), ("P", C.c_double), ("V", C.c_double),
                ("ld", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("ox", C.c_int),
                ("oy", C.c_int), ("step", C.c_int)]


assert C.sizeof(SgWindow) == 96


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights (output layer scaled down), zero biases; the
    layout of include/lpbf_surrogate.h."""
    rng = np.random.default_rng(seed)
    out = []
    for q, (ci, co) in enumerate([(4, CH), (CH, CH), (CH, CH), (CH, CH), (CH, CH), (CH, 1)]):
        W = rng.normal(0.0, math.sqrt(2.0 / (9 * ci)), (9, ci, co))
        if q == 5:
            W *= 0.1
        out += [W.ravel(), np.zeros(co)]
    p = np.concatenate(out).astype(np.float32)
    assert p.size == NPARAM
    return p


def window_origin(x_l, y_l, dx, dy, nx, ny):
    """Window of WIN x WIN cells centred on the laser, clamped to the plate."""
    ox = int(round(x_l / dx)) - WIN // 2
    oy = int(round(y_l / dy)) - WIN // 2
    return min(max(ox, 0), nx - WIN), min(max(oy, 0), ny - WIN)


class Surrogate:
    """Parameters, Adam state and workspaces on the GPU."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0, device="cuda"):
        import torch
        self.torch = torch
        self.dev = device
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.from_numpy(p.copy()).to(device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.t = 0
        self.pack = torch.zeros(lib().lpbf_sg_pack_bytes(), dtype=torch.uint8, device=device)
        self.loss = torch.zeros(1, dtype=torch.float64, device=device)
        self._work = {}
        self.repack()

    def _stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def _ws(self, batch, training):
        key = (batch, training)
        if key not in self._work:
            n = lib().lpbf_sg_work_bytes(batch, int(training))
            if n == 0:
                raise LpbfError(f"batch {batch} outside the surrogate's range")
            self._work[key] = self.torch.empty(n, dtype=self.torch.uint8, device=self.dev)
        return self._work[key]

    def repack(self):
        lib().lpbf_sg_pack(self.params.data_ptr(), self.pack.data_ptr(), self._stream())

    def forward(self, x0, out=None):
        """x0: bf16 [B, 128, 128, 32] -> FP32 [B, 128, 128]."""
        B = x0.shape[0]
        y = out if out is not None else self.torch.empty((B, WIN, WIN), dtype=self.torch.float32,
                                                           device=self.dev)
        lib().lpbf_sg_forward(self.pack.data_ptr(), x0.data_ptr(), B,
                              self._ws(B, False).data_ptr(), y.data_ptr(), self._stream())
        return y

    def loss_grad(self, x0, target, y=None):
        B = x0.shape[0]
        lib().lpbf_sg_loss_grad(self.pack.data_ptr(), x0.data_ptr(), target.data_ptr(), B,
                                self._ws(B, True).data_ptr(), self.grad.data_ptr(),
                                self.loss.data_ptr(), y.data_ptr() if y is not None else None,
                                self._stream())
        return self.loss, self.grad

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        lib().lpbf_sg_adam(self.params.data_ptr(), self.grad.data_ptr(), self.m.data_ptr(),
                           self.v.data_ptr(), NPARAM, self.t, lr, beta1, beta2, eps,
                           self._stream())
        self.repack()

    def train_step(self, x0, target, lr=1e-3):
        self.loss_grad(x0, target)
        self.adam(lr)
        return self.loss


def windows_buffer(wins: list[SgWindow], device="cuda"):
    import torch
    raw = bytes((SgWindow * len(wins))(*wins))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


def gather(win_dev, n, x0, T0):
    lib().lpbf_sg_gather(win_dev.data_ptr(), n, x0.data_ptr(), T0.data_ptr(),
                         _cur_stream())


def target(win_dev, n, T0, tgt):
    lib().lpbf_sg_target(win_dev.data_ptr(), n, T0.data_ptr(), tgt.data_ptr(), _cur_stream())


def apply(win_dev, n, T0, y, overlap=8, skip=None, maxdiff=None):
    lib().lpbf_sg_apply(win_dev.data_ptr(), n, T0.data_ptr(), y.data_ptr(), overlap,
                        skip.data_ptr() if skip is not None else None,
                        maxdiff.data_ptr() if maxdiff is not None else None, _cur_stream())


def _cur_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ---- training data from FD runs ------------------------------------------------------

@dataclass
class DataSpec:
    n_runs: int = 64
    nx: int = 640
    ny: int = 384
    dx: float = 5e-6
    dt: float = 1.2e-6
    steps: int = 4096                 # per run; one pair every 16 steps
    P: tuple = (200.0, 350.0)
    V: tuple = (0.8, 1.6)
    r0: float = 50e-6
    n_tracks: int = 3
    hatch: float = 80e-6
    seed: int = 0


def plate_material():
    from . import Material
    return Material(k=7.0, rho=4430.0, cp=526.0, Tm=1609.0, eps=0.40, eta=0.35)


def make_runs(spec: DataSpec):
    """Per run: (StepConst, plan[steps,3], P, V) for a raster over the plate."""
    from . import Domain, Process, centred_raster, make_const, raster_plan
    rng = np.random.default_rng(spec.seed)
    mat = plate_material()
    runs = []
    for r in range(spec.n_runs):
        P = float(rng.uniform(*spec.P))
        V = float(rng.uniform(*spec.V))
        pr = Process(P=P, V=V, r0=spec.r0, h=50e-6, Tamb=300.0, Tsub=300.0)
        dom = Domain(Lx=(spec.nx - 1) * spec.dx, Ly=(spec.ny - 1) * spec.dx, nx=spec.nx,
                     ny=spec.ny)
        # enough tracks that each stays within 70% of the plate width
        n_tracks = max(spec.n_tracks, math.ceil(V * spec.dt * spec.steps / (0.7 * dom.Lx)))
        rs = centred_raster(dom.Lx, dom.Ly, V, spec.dt, spec.steps, n_tracks, spec.hatch)
        plan = raster_plan(rs["x0"], rs["y0"], rs["length"], rs["hatch"], rs["n_tracks"],
                           rs["V"], spec.dt, spec.steps)
        runs.append((make_const(mat, pr, dom, spec.dt), plan, P, V))
    return runs


def generate(spec: DataSpec, device="cuda", log=None):
    """Runs the FD solver (FP32 batch kernel) on spec.n_runs plates and returns
    the pairs: x (bf16 [N,128,128,4]), T0 (FP32 theta at t), target (FP32
    (T_{t+16}-T_t)/3000); windows centred on the laser at t."""
    import torch

    from .batch import F32Batch, InstanceSpec
    runs = make_runs(spec)
    specs = [InstanceSpec(k=k, plan=plan, T0=300.0, T_ref=300.0, Tm=1609.0, size_class=0,
                          stats=False) for k, plan, _, _ in runs]
    b = F32Batch(specs, bc=LPBF_BC_MIRROR, device=device)
    b.set_tiling(0, 4, 0)
    n_blocks = spec.steps // STEPS - 1
    N = n_blocks * spec.n_runs
    xs = torch.empty((N, WIN, WIN, 4), dtype=torch.bfloat16, device=device)
    T0s = torch.empty((N, WIN, WIN), dtype=torch.float32, device=device)
    tg = torch.empty((N, WIN, WIN), dtype=torch.float32, device=device)
    x0 = torch.empty((spec.n_runs, WIN, WIN, CH), dtype=torch.bfloat16, device=device)
    T0 = torch.empty((spec.n_runs, WIN, WIN), dtype=torch.float32, device=device)
    t0 = time.perf_counter()
    for blk in range(n_blocks):
        s = blk * STEPS
        wins = []
        for i, (k, plan, P, V) in enumerate(runs):
            xl, yl, q = plan[s]
            ox, oy = window_origin(xl, yl, k.dx, k.dy, k.nx, k.ny)
            th = b.theta(i)
            wins.append(SgWindow(th.data_ptr(), th.data_ptr(), b._bufs[i][2].data_ptr(), 300.0,
                                 k.dx, k.dy, k.inv_r0_sq, P, V, th.shape[1], k.nx, k.ny, ox, oy,
                                 s))
        wd = windows_buffer(wins, device)
        gather(wd, len(wins), x0, T0)
        b.enqueue(s, STEPS)
        # the fields moved to the other buffer: point the windows at the new state
        wins2 = []
        for i, w in enumerate(wins):
            th = b.theta(i)
            w2 = SgWindow.from_buffer_copy(bytes(w))
            w2.theta = th.data_ptr()
            w2.theta_out = th.data_ptr()
            wins2.append(w2)
        wd2 = windows_buffer(wins2, device)
        sl = slice(blk * spec.n_runs, (blk + 1) * spec.n_runs)
        target(wd2, len(wins2), T0, tg[sl])
        xs[sl] = x0[..., :4]
        T0s[sl] = T0
    torch.cuda.synchronize()
    check()
    if log:
        log(f"[surrogate] {N} pairs from {spec.n_runs} FD runs of {spec.steps} steps in "
            f"{time.perf_counter() - t0:.1f} s")
    b.close()
    return xs, T0s, tg


def expand(x4, idx, out):
    """x0 [B,128,128,32] bf16 from stored 4-channel inputs (channels 4..31 = 0)."""
    out.zero_()
    out[..., :4] = x4[idx]
    return out


def train(model: Surrogate, xs, tg, iters: int, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, log=None, log_every: int = 100, time_budget: float | None = None):
    import torch
    N = xs.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    x0 = torch.empty((batch, WIN, WIN, CH), dtype=torch.bfloat16, device=xs.device)
    perm = torch.randperm(N, generator=g)
    pos = 0
    losses = []
    t0 = time.perf_counter()
    for it in range(iters):
        if pos + batch > N:
            perm = torch.randperm(N, generator=g)
            pos = 0
        idx = perm[pos:pos + batch].to(xs.device)
        pos += batch
        expand(xs, idx, x0)
        loss = model.train_step(x0, tg[idx].contiguous(), lr)
        if (it + 1) % log_every == 0 or it == iters - 1:
            lv = float(loss.item())
            losses.append((it + 1, lv))
            if log:
                log(f"[surrogate] iter {it + 1}: loss {lv:.3e} "
                    f"({(time.perf_counter() - t0) / (it + 1) * 1e3:.2f} ms/iter)")
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            break
    torch.cuda.synchronize()
    check()
    return losses


def evaluate(model: Surrogate, xs, T0s, tg, idx, batch: int = 64):
    """Errors in kelvin of the trained network and of persistence (T_{t+16} =
    T_t, the method without the network) on windows idx: (rms, max) each."""
    import torch
    x0 = torch.empty((batch, WIN, WIN, CH), dtype=torch.bfloat16, device=xs.device)
    se_n = se_p = 0.0
    mx_n = mx_p = 0.0
    cnt = 0
    for s in range(0, len(idx), batch):
        j = idx[s:s + batch]
        if len(j) < batch:
            break
        jd = torch.as_tensor(j, device=xs.device)
        expand(xs, jd, x0)
        y = model.forward(x0)
        e_n = (y - tg[jd]) * 3000.0
        e_p = tg[jd] * 3000.0
        se_n += float((e_n.double() ** 2).sum())
        se_p += float((e_p.double() ** 2).sum())
        mx_n = max(mx_n, float(e_n.abs().max()))
        mx_p = max(mx_p, float(e_p.abs().max()))
        cnt += e_n.numel()
    return {"rms_net": math.sqrt(se_n / cnt), "max_net": mx_n,
            "rms_persistence": math.sqrt(se_p / cnt), "max_persistence": mx_p, "windows": cnt // WIN // WIN}


# ---- surrogate-accelerated run ----------------------------------------------------------

@dataclass
class RunSpec:
    nx: int = 4096
    ny: int = 4096
    dx: float = 5e-6
    dt: float = 1.2e-6
    steps: int = 40000
    P: float = 300.0
    V: float = 1.0
    r0: float = 50e-6
    n_tracks: int = 10
    hatch: float = 80e-6
    depth: int = 4
    rows: int = 0
    recheck_every: int = 8
    tol_K: float = 2.0
    overlap: int = 8


def run_setup(rs: RunSpec):
    from . import Domain, Process, centred_raster, make_const, raster_plan
    mat = plate_material()
    pr = Process(P=rs.P, V=rs.V, r0=rs.r0, h=50e-6, Tamb=300.0, Tsub=300.0)
    dom = Domain(Lx=(rs.nx - 1) * rs.dx, Ly=(rs.ny - 1) * rs.dx, nx=rs.nx, ny=rs.ny)
    r = centred_raster(dom.Lx, dom.Ly, rs.V, rs.dt, rs.steps, rs.n_tracks, rs.hatch)
    plan = raster_plan(r["x0"], r["y0"], r["length"], r["hatch"], r["n_tracks"], r["V"],
                       rs.dt, rs.steps + STEPS)
    return mat, pr, dom, make_const(mat, pr, dom, rs.dt), plan


class _F32Field:
    """Two pitched FP32 theta buffers, the plan on the device and the step
    workspace (theta = T - 300 K, field starts at 300 K)."""

    def __init__(self, k, plan, depth, rows, device="cuda"):
        import torch
        self.k, self.depth, self.rows = k, depth, rows
        self.ld = (k.nx + 3) // 4 * 4
        self.a = torch.zeros((k.ny, self.ld), dtype=torch.float32, device=device)
        self.b = torch.zeros_like(self.a)
        self.plan = torch.from_numpy(np.ascontiguousarray(plan)).to(device)
        self.work = torch.empty(lib().lpbf_f32_workspace_bytes(k.nx, k.ny, depth, rows) // 8 + 8,
                                dtype=torch.float64, device=device)

    def advance(self, s0, n, stream):
        s = s0
        while s < s0 + n:
            d = min(self.depth, s0 + n - s)
            lib().lpbf_launch_step_f32(self.a.data_ptr(), self.b.data_ptr(), self.ld, self.k,
                                       300.0, 0, 300.0, 1609.0, self.plan.data_ptr(), s, d,
                                       self.rows, None, self.work.data_ptr(), stream)
            self.a, self.b = self.b, self.a
            s += d

    def T(self):
        return self.a[:, :self.k.nx].double() + 300.0


def fp64_reference(k, plan, steps, device="cuda"):
    """The FP64 parity path (lpbf_launch_step_ex, mirror edges) over the run;
    its step equals the CPU oracle's to 1e-9 K (tests/test_gpu_f64.py)."""
    import torch

    from ._lib import StepConst as SC
    a = torch.full((k.ny, k.nx), 300.0, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    kk = SC.from_buffer_copy(bytes(k))
    st = torch.cuda.current_stream().cuda_stream
    for s in range(steps):
        kk.x_l, kk.y_l, kk.q_scale = plan[s]
        lib().lpbf_launch_step_ex(a.data_ptr(), b.data_ptr(), kk, 0, 0, 300.0, 1609.0, None, st)
        a, b = b, a
    torch.cuda.synchronize()
    check()
    return a


def accelerated_run(model: Surrogate | None, rs: RunSpec, log=None):
    """Plain FP32 FD over the whole plate, 16 steps at a time; with a model,
    after every 16 steps the 128 x 128 window centred on the laser at the start
    of the block is replaced by the surrogate's prediction from the field at
    the start (blended over rs.overlap border cells).  Every rs.recheck_every
    calls the prediction is compared with the FD result of the same window; a
    difference above rs.tol_K keeps the FD result (fallback) and the next
    rs.recheck_every - 1 calls are also left to FD.  Returns the final FP32
    field (as T), wall seconds and counters."""
    import torch
    mat, pr, dom, k, plan = run_setup(rs)
    f = _F32Field(k, plan, rs.depth, rs.rows)
    stream = torch.cuda.current_stream().cuda_stream
    x0 = torch.empty((1, WIN, WIN, CH), dtype=torch.bfloat16, device="cuda")
    T0 = torch.empty((1, WIN, WIN), dtype=torch.float32, device="cuda")
    y = torch.empty((1, WIN, WIN), dtype=torch.float32, device="cuda")
    md = torch.zeros(1, dtype=torch.float32, device="cuda")
    skip = torch.zeros(1, dtype=torch.int32, device="cuda")
    calls = rechecks = fallbacks = applied = 0
    off_until = -1
    since_check = 0     # calls applied since the last recheck; 0: the next call is checked
    diffs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_blocks = rs.steps // STEPS
    for blk in range(n_blocks):
        s = blk * STEPS
        use = model is not None and blk > off_until and plan[s][2] != 0.0
        if use:
            ox, oy = window_origin(plan[s][0], plan[s][1], k.dx, k.dy, k.nx, k.ny)
            w = SgWindow(f.a.data_ptr(), 0, f.plan.data_ptr(), 300.0, k.dx, k.dy, k.inv_r0_sq,
                         rs.P, rs.V, f.ld, k.nx, k.ny, ox, oy, s)
            wd = windows_buffer([w])
            gather(wd, 1, x0, T0)
            model.forward(x0, out=y)
        f.advance(s, STEPS, stream)
        if use:
            calls += 1
            w.theta_out = f.a.data_ptr()
            wd = windows_buffer([w])
            if since_check == 0:
                # compare with the FD result of the same window, write nothing
                md.zero_()
                skip.fill_(1)
                apply(wd, 1, T0, y, rs.overlap, skip, md)
                d = float(md.item())
                diffs.append(d)
                rechecks += 1
                if d > rs.tol_K:
                    # keep FD here and for the next recheck_every - 1 blocks;
                    # the first call after that is checked again
                    fallbacks += 1
                    off_until = blk + rs.recheck_every - 1
                    continue
            apply(wd, 1, T0, y, rs.overlap, None, None)
            applied += 1
            since_check = (since_check + 1) % rs.recheck_every
    f.advance(n_blocks * STEPS, rs.steps - n_blocks * STEPS, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    check()
    out = {"wall_s": wall, "sim_s": rs.steps * rs.dt, "calls": calls, "applied": applied,
           "rechecks": rechecks, "fallbacks": fallbacks,
           "fallback_rate": fallbacks / rechecks if rechecks else None,
           "recheck_maxdiff_K": diffs}
    if log:
        log(f"[surrogate-run] {'model' if model is not None else 'plain'}: {wall:.2f} s wall, "
            f"{calls} calls, {applied} applied, {fallbacks}/{rechecks} rechecks fell back")
    return f.T(), out
