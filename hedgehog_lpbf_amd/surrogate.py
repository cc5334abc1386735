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

from ._lib import (LPBF_BC_MIRROR, LaserTick, LpbfError, MrConfig, MrCounters,  # noqa: F401
                   MrWindow, StepConst, check, lib)

WIN = 128
CH = 32
NPARAM = 38465
STEPS = 16


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


def _cur_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


# ---- the two-level run's grids (include/lpbf_multirate.h) ----------------------------

def coarse_const(k: StepConst) -> StepConst:
    """The coarse grid's constants, as lpbf_mr_create derives them: 2x2 cells,
    four fine steps per coarse step, the same physics."""
    c = StepConst.from_buffer_copy(bytes(k))
    c.nx, c.ny = k.nx // 2, k.ny // 2
    c.dx, c.dy = 2.0 * k.dx, 2.0 * k.dy
    c.inv_dx2, c.inv_dy2 = 1.0 / (c.dx * c.dx), 1.0 / (c.dy * c.dy)
    c.dt = 4.0 * k.dt
    return c


def coarse_plan(plan: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """One coarse tick per four fine ticks, as lpbf_mr_create builds them: the
    mean beam strength at its strength-weighted mean position, shifted by half
    a fine cell (coarse cell u is centred on fine position 2u + 1/2 while the
    step kernel places it at 2u)."""
    n4 = plan.shape[0] // 4
    p = plan[:4 * n4].reshape(n4, 4, 3)
    q = p[..., 2]
    qs = q.sum(axis=1)
    safe = np.where(qs > 0, qs, 1.0)
    x = np.where(qs > 0, (q * p[..., 0]).sum(axis=1) / safe, p[..., 0].mean(axis=1))
    y = np.where(qs > 0, (q * p[..., 1]).sum(axis=1) / safe, p[..., 1].mean(axis=1))
    return np.stack([x - 0.5 * dx, y - 0.5 * dy, 0.25 * qs], axis=1)


def window_origins(plan: np.ndarray, n_blocks: int, nx: int, ny: int, dx: float, dy: float):
    """Window origins of blocks 0..n_blocks (lpbf_mr_window_origins: the laser
    at 3/4 of the window along its travel, centred across it, even, inside
    the plate)."""
    p = np.ascontiguousarray(plan[:STEPS * n_blocks + 1], dtype=np.float64)
    ox = np.zeros(n_blocks + 1, dtype=np.int32)
    oy = np.zeros(n_blocks + 1, dtype=np.int32)
    rc = lib().lpbf_mr_window_origins(p.ctypes.data, n_blocks, nx, ny, dx, dy, ox.ctypes.data,
                                      oy.ctypes.data)
    if rc:
        check()
        raise LpbfError(f"lpbf_mr_window_origins failed: {rc}")
    return ox, oy


def _mr_windows(wins, device="cuda"):
    import torch
    raw = bytes((MrWindow * len(wins))(*wins))
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


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
    """Training pairs of the two-level run's network from fine FP32 FD runs
    (the batch kernel) and their coarse companions: every 16 steps, each
    run's fine field is restricted onto its coarse grid, the coarse grid takes
    its four steps, and the window centred on the laser yields the network
    input (lpbf_mr_gather) and, after the fine run's 16 steps, the target
    (fine increment - prolonged coarse increment) / 100 K (lpbf_mr_target).
    Returns x (bf16 [N,128,128,4]), D (FP32 [N,128,128], prolonged coarse
    increment, K) and the target (FP32 [N,128,128])."""
    import torch

    from .batch import F32Batch, InstanceSpec
    if spec.nx % 2 or spec.ny % 2:
        raise ValueError("the coarse grid needs even plate sides")
    runs = make_runs(spec)
    specs = [InstanceSpec(k=k, plan=plan, T0=300.0, T_ref=300.0, Tm=1609.0, size_class=0,
                          stats=False) for k, plan, _, _ in runs]
    cspecs = [InstanceSpec(k=coarse_const(k), plan=coarse_plan(plan, k.dx, k.dy), T0=300.0,
                           T_ref=300.0, Tm=1609.0, size_class=0, stats=False)
              for k, plan, _, _ in runs]
    b = F32Batch(specs, bc=LPBF_BC_MIRROR, device=device)
    cb = F32Batch(cspecs, bc=LPBF_BC_MIRROR, device=device)
    b.set_tiling(0, 4, 0)
    cb.set_tiling(0, 2, 0)            # 4 coarse steps: two launches, the field stays put
    n = spec.n_runs
    nxc, nyc = spec.nx // 2, spec.ny // 2
    ldc = cb.theta(0).shape[1]
    c0 = torch.empty((n, nyc, ldc), dtype=torch.float32, device=device)
    n_blocks = spec.steps // STEPS - 1
    N = n_blocks * n
    xs = torch.empty((N, WIN, WIN, 4), dtype=torch.bfloat16, device=device)
    Ds = torch.empty((N, WIN, WIN), dtype=torch.float32, device=device)
    tg = torch.empty((N, WIN, WIN), dtype=torch.float32, device=device)
    x0 = torch.empty((n, WIN, WIN, CH), dtype=torch.bfloat16, device=device)
    T0 = torch.empty((n, WIN, WIN), dtype=torch.float32, device=device)
    plans = [b._bufs[i][2] for i in range(n)]
    origins = [window_origins(plan, n_blocks, k.nx, k.ny, k.dx, k.dy) for k, plan, _, _ in runs]
    st = _cur_stream()
    t0 = time.perf_counter()
    for blk in range(n_blocks):
        s = blk * STEPS
        wins = []
        for i, (k, plan, _, _) in enumerate(runs):
            th, cth = b.theta(i), cb.theta(i)
            lib().lpbf_mr_restrict(th.data_ptr(), th.shape[1], k.nx, k.ny, cth.data_ptr(), ldc, st)
            lib().lpbf_mr_restrict(th.data_ptr(), th.shape[1], k.nx, k.ny, c0[i].data_ptr(), ldc,
                                   st)
        cb.enqueue(4 * blk, 4)
        for i, (k, plan, _, _) in enumerate(runs):
            th = b.theta(i)
            ox, oy = int(origins[i][0][blk]), int(origins[i][1][blk])
            wins.append(MrWindow(th.data_ptr(), c0[i].data_ptr(), cb.theta(i).data_ptr(),
                                 plans[i].data_ptr(), k.dx, k.dy, k.inv_r0_sq, th.shape[1], ox,
                                 oy, ox, oy, ldc, nxc, nyc, s))
        wd = _mr_windows(wins, device)
        sl = slice(blk * n, (blk + 1) * n)
        lib().lpbf_mr_gather(wd.data_ptr(), n, x0.data_ptr(), Ds[sl].data_ptr(), T0.data_ptr(), st)
        b.enqueue(s, STEPS)           # four launches: the field is back in the same buffer
        lib().lpbf_mr_target(wd.data_ptr(), n, T0.data_ptr(), Ds[sl].data_ptr(), tg[sl].data_ptr(),
                             st)
        xs[sl] = x0[..., :4]
    torch.cuda.synchronize()
    check()
    if log:
        log(f"[surrogate] {N} pairs from {spec.n_runs} FD runs of {spec.steps} steps in "
            f"{time.perf_counter() - t0:.1f} s")
    b.close()
    cb.close()
    return xs, Ds, tg


def expand(x4, idx, out):
    """x0 [B,128,128,32] bf16 from stored 4-channel inputs (channels 4..31 = 0)."""
    out.zero_()
    out[..., :4] = x4[idx]
    return out


def train(model: Surrogate, xs, tg, iters: int, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, log=None, log_every: int = 100, time_budget: float | None = None,
          lr_final: float | None = None):
    """Adam on random batches; with lr_final the learning rate decays
    geometrically from lr to lr_final over the iterations."""
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
        lr_it = lr if lr_final is None else lr * (lr_final / lr) ** (it / max(iters - 1, 1))
        loss = model.train_step(x0, tg[idx].contiguous(), lr_it)
        if (it + 1) % log_every == 0 or it == iters - 1:
            lv = float(loss.item())
            losses.append((it + 1, lv))
            if log:
                log(f"[surrogate] iter {it + 1}: loss {lv:.3e} lr {lr_it:.2e} "
                    f"({(time.perf_counter() - t0) / (it + 1) * 1e3:.2f} ms/iter)")
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            break
    torch.cuda.synchronize()
    check()
    return losses


def evaluate(model: Surrogate, xs, tg, idx, batch: int = 64, overlap: int = 8):
    """Errors in kelvin on windows idx, over the cells at least `overlap` from
    the window edge (where the run applies the network at full weight): the
    trained network's residual error and that of the two-level step without
    the network (the prolonged coarse increment alone, residual 0): rms and
    max each."""
    import torch
    x0 = torch.empty((batch, WIN, WIN, CH), dtype=torch.bfloat16, device=xs.device)
    inner = (slice(None), slice(overlap, WIN - overlap), slice(overlap, WIN - overlap))
    se_n = se_b = 0.0
    mx_n = mx_b = 0.0
    cnt = 0
    for s in range(0, len(idx), batch):
        j = idx[s:s + batch]
        if len(j) < batch:
            break
        jd = torch.as_tensor(j, device=xs.device)
        expand(xs, jd, x0)
        y = model.forward(x0)
        e_n = ((y - tg[jd]) * 100.0)[inner]
        e_b = (tg[jd] * 100.0)[inner]
        se_n += float((e_n.double() ** 2).sum())
        se_b += float((e_b.double() ** 2).sum())
        mx_n = max(mx_n, float(e_n.abs().max()))
        mx_b = max(mx_b, float(e_b.abs().max()))
        cnt += e_n.numel()
    return {"rms_net": math.sqrt(se_n / cnt), "max_net": mx_n,
            "rms_without_net": math.sqrt(se_b / cnt), "max_without_net": mx_b,
            "windows": cnt // (WIN - 2 * overlap) ** 2}


def save_params(model: Surrogate, path: str):
    np.save(path, model.params.cpu().numpy())


def load_params(path: str, device="cuda") -> Surrogate:
    return Surrogate(params=np.load(path, allow_pickle=False), device=device)


# ---- runs: plain FP32, FP64 reference, two-level ----------------------------------------

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


def plain_run(rs: RunSpec, log=None):
    """The plain FP32 FD run of the whole plate (the FP32 step kernel with the
    autotuner's depth and band height, stats off).  Returns (T as FP64 on the
    device, wall seconds)."""
    import torch
    mat, pr, dom, k, plan = run_setup(rs)
    st = _cur_stream()
    pdev = torch.from_numpy(np.ascontiguousarray(plan)).cuda()
    bd, br, bms = C.c_int(), C.c_int(), C.c_double()
    rc = lib().lpbf_f32_autotune(k, 300.0, LPBF_BC_MIRROR, 300.0, 1609.0, pdev.data_ptr(),
                                 plan.shape[0], 4, st, C.byref(bd), C.byref(br), C.byref(bms))
    if rc:
        check()
        raise LpbfError(f"autotune failed: {rc}")
    depth, rows = bd.value, br.value
    ld = (k.nx + 3) // 4 * 4
    a = torch.zeros((k.ny, ld), dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    work = torch.empty(lib().lpbf_f32_workspace_bytes(k.nx, k.ny, depth, rows) // 8 + 8,
                       dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = 0
    while s < rs.steps:
        d = min(depth, rs.steps - s)
        lib().lpbf_launch_step_f32(a.data_ptr(), b.data_ptr(), ld, k, 300.0, LPBF_BC_MIRROR,
                                   300.0, 1609.0, pdev.data_ptr(), s, d, rows, None,
                                   work.data_ptr(), st)
        a, b = b, a
        s += d
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    check()
    if log:
        log(f"[plain] {rs.nx}x{rs.ny} {rs.steps} steps: {wall:.2f} s (depth {depth}, "
            f"rows {rows})")
    return a[:, :k.nx].double() + 300.0, wall


def fp64_reference(k, plan, steps, device="cuda"):
    """The FP64 parity path over the run (the temporally blocked FP64 kernel at
    depth 2, bitwise the single-step field, which equals the CPU oracle's to
    1e-9 K), mirror edges."""
    import torch
    a = torch.full((k.ny, k.nx), 300.0, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    pdev = torch.from_numpy(np.ascontiguousarray(plan)).to(device)
    st = _cur_stream()
    s = 0
    while s < steps:
        d = min(2, steps - s)
        lib().lpbf_launch_step_f64_blocked(a.data_ptr(), b.data_ptr(), k, 0, LPBF_BC_MIRROR,
                                           300.0, 1609.0, pdev.data_ptr(), s, d, 0, None, None,
                                           st)
        a, b = b, a
        s += d
    torch.cuda.synchronize()
    check()
    return a


class TwoLevelRun:
    """lpbf_mr_* run object: coarse far field plus the fine melt-pool window,
    advanced by `model` (with FD rechecks and fallback) or, without a model,
    by the FD patch every block."""

    def __init__(self, rs: RunSpec, model: Surrogate | None = None, tol_K: float | None = None,
                 recheck_every: int | None = None):
        import torch
        self.torch = torch
        self.rs = rs
        self.mat, self.pr, self.dom, self.k, self.plan = run_setup(rs)
        cfg = MrConfig()
        cfg.k = self.k
        cfg.T_ref = 300.0
        cfg.tol_K = rs.tol_K if tol_K is None else tol_K
        cfg.n_steps = rs.steps
        cfg.recheck_every = rs.recheck_every if recheck_every is None else recheck_every
        cfg.overlap = rs.overlap
        cfg.use_network = int(model is not None)
        cfg.coarse_depth = 0                  # autotuned
        cfg.coarse_rows = 0
        self.model = model
        plan = np.ascontiguousarray(self.plan, dtype=np.float64)
        h = C.c_void_p()
        rc = lib().lpbf_mr_create(C.byref(cfg), plan.ctypes.data,
                                  model.pack.data_ptr() if model is not None else None,
                                  _cur_stream(), C.byref(h))
        if rc:
            check()
            raise LpbfError(f"lpbf_mr_create failed: {rc}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lpbf_mr_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def advance(self, n_blocks: int):
        rc = lib().lpbf_mr_advance(self._h, n_blocks)
        if rc:
            check()
            raise LpbfError(f"lpbf_mr_advance failed: {rc}")

    def field(self):
        T = self.torch.empty((self.k.ny, self.k.nx), dtype=self.torch.float64, device="cuda")
        rc = lib().lpbf_mr_field(self._h, T.data_ptr())
        if rc:
            raise LpbfError(f"lpbf_mr_field failed: {rc}")
        return T

    def counters(self) -> dict:
        c = MrCounters()
        lib().lpbf_mr_get_counters(self._h, C.byref(c))
        return {f: getattr(c, f) for f, _ in MrCounters._fields_}

    def recheck_diffs(self):
        n = lib().lpbf_mr_recheck_diffs(self._h, None, 0)
        out = np.zeros(max(n, 1), dtype=np.float32)
        lib().lpbf_mr_recheck_diffs(self._h, out.ctypes.data, n)
        return out[:n]


def two_level_run(rs: RunSpec, model: Surrogate | None = None, log=None, **kw):
    """Whole run timed: returns (T as FP64 on the device, wall seconds,
    counters, recheck differences)."""
    import torch
    r = TwoLevelRun(rs, model, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r.advance(rs.steps // STEPS)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    check()
    T = r.field()
    torch.cuda.synchronize()
    cnt, diffs = r.counters(), r.recheck_diffs()
    r.close()
    if log:
        log(f"[two-level] {'network' if model is not None else 'FD window'}: {wall:.2f} s, "
            f"{cnt['calls']} calls, {cnt['applied']} applied, {cnt['fallbacks']}/"
            f"{cnt['rechecks']} rechecks fell back, {cnt['fd_blocks']} FD blocks")
    return T, wall, cnt, diffs
