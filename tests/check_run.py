"""CPU-versus-GPU deviation report for one run, the counterpart of the
reference's ``hedgehog_lpbf_gpu --check`` (src/gpu/lpbf_cuda.cu:485-514).

It runs the CPU oracle and the device-resident GPU run (lpbf_run_* through
hedgehog_lpbf_amd.DeviceRun) over the same steps and prints the final-field
max|dT| and max relative dT (relative to max(|T_cpu|, 1)), and the CPU, GPU
and absolute differences of peak T, peak W and the cruise-gated steady
sub-grid width.  It passes when ``max_rel <= rtol or max_abs <= atol`` and
exits 1 on FAIL, 3 when no GPU is present, as the reference does
(lpbf_cuda.cu:467-472, 509-513).  Defaults are the reference's rtol 1e-6 and
atol 1e-3 K (src/common.h:96-97).

The comparison lives with the tests because the CPU side is the oracle; the
lpbf_gpu binary carries no CPU solver.

    python -m tests.check_run --case=3.1 --physics=ext --kmult=2.5 --tend=2e-4
"""
from __future__ import annotations

import argparse
import os
import sys
from dataclasses import dataclass

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


@dataclass
class RunResult:
    """Reference RunResult (lpbf_cuda.cu:300-304): run maxima and final field."""
    peak_T: float
    peak_W: float
    steady_W_sub: float
    T: np.ndarray


@dataclass
class Report:
    max_abs: float
    max_rel: float
    d_peak_T: float
    d_peak_W: float
    d_steady: float
    passed: bool


def compare(cpu: RunResult, gpu: RunResult, rtol: float, atol: float) -> Report:
    """The reference's deviation measures and pass rule (lpbf_cuda.cu:488-509)."""
    d = np.abs(gpu.T - cpu.T)
    max_abs = float(d.max()) if d.size else 0.0
    max_rel = float((d / np.maximum(np.abs(cpu.T), 1.0)).max()) if d.size else 0.0
    return Report(max_abs, max_rel, abs(cpu.peak_T - gpu.peak_T), abs(cpu.peak_W - gpu.peak_W),
                  abs(cpu.steady_W_sub - gpu.steady_W_sub),
                  (max_rel <= rtol) or (max_abs <= atol))


def parse(argv):
    """The reference CLI's run options (src/common.h:100-146) plus --precision."""
    ap = argparse.ArgumentParser(prog="check_run", description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", default="0")
    ap.add_argument("--physics", default="base", choices=("base", "ext"))
    ap.add_argument("--tend", type=float, default=1.5e-3)
    for o in ("P", "V", "r0", "eta", "dsub", "hconv"):
        ap.add_argument(f"--{o}", type=float, default=-1.0)
    ap.add_argument("--kmult", type=float, default=1.0)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--check-rtol", dest="rtol", type=float, default=1e-6)
    ap.add_argument("--check-atol", dest="atol", type=float, default=1e-3)
    ap.add_argument("--precision", type=int, default=64, choices=(64, 32))
    ap.add_argument("--batch", type=int, default=64)
    return ap.parse_args(argv)


def setup(a):
    """Case, overrides and dt as the reference main (lpbf_cuda.cu:453-466)."""
    import hedgehog_lpbf_amd as h
    mat, pr, dom = h.Material(), h.Process(), h.Domain()
    h.apply_case(a.case, pr)
    if a.P > 0:
        pr.P = a.P
    if a.V > 0:
        pr.V = a.V
    if a.r0 > 0:
        pr.r0 = a.r0
    if a.eta > 0:
        mat.eta = a.eta
    if a.dsub > 0:
        pr.dSub = a.dsub
    if a.hconv >= 0:
        pr.hConv = a.hconv
    phys = 1 if a.physics == "ext" else 0
    dt = h.dt_cfl_ext(mat, dom, a.kmult) if phys else h.dt_cfl(mat, dom)
    return mat, pr, dom, phys, dt, int(a.tend / dt)


def run_cpu(a) -> RunResult:
    """The oracle's full run (oracle/lpbf_oracle.c ora_run), the reference's
    run_cpu (lpbf_cuda.cu:423-448)."""
    from oracle import pyoracle as ora
    mat, pr, dom, phys, dt, n = setup(a)
    s = ora.defaults(None)
    for f in ("k", "rho", "cp", "Tm", "eps", "eta"):
        setattr(s.mat, f, getattr(mat, f))
    for f in ("P", "V", "r0", "h", "hConv", "dSub", "Tamb", "Tsub", "radiation"):
        setattr(s.proc, f, getattr(pr, f))
    T, _, _, _, summ = ora.run(s, dt, n, phys=phys, kmult=a.kmult)
    return RunResult(summ[0], summ[1], summ[2], T)


def run_gpu(a) -> RunResult:
    import hedgehog_lpbf_amd as h
    mat, pr, dom, phys, dt, n = setup(a)
    k = h.make_const(mat, pr, dom, dt)
    k.kmult = a.kmult
    plan, x_csv, wrapped = h.single_track_plan(pr.V, dt, dom.Lx, dom.Ly, n)
    with h.DeviceRun(k, Tm=mat.Tm, Lx=dom.Lx, Ly=dom.Ly, Tamb=pr.Tamb, plan=plan, x_csv=x_csv,
                     wrapped=wrapped, phys=phys, precision=a.precision, batch=a.batch,
                     keep_records=False) as r:
        r.advance(n)
        T = r.download()
        peak_T, peak_W, steady, _ = r.summary()
    return RunResult(peak_T, peak_W, steady, T)


def main(argv=None) -> int:
    a = parse(sys.argv[1:] if argv is None else argv)
    import torch
    ndev = torch.cuda.device_count()
    if ndev == 0:
        print("No HIP device found.", file=sys.stderr)
        return 3
    if not 0 <= a.device < ndev:
        print(f"Invalid --device={a.device} (have {ndev})", file=sys.stderr)
        return 3
    torch.cuda.set_device(a.device)
    *_, n = setup(a)
    print(f"[--check] running CPU and GPU for {n} steps (case {a.case}, {a.physics}, "
          f"FP{a.precision})...", flush=True)
    cpu = run_cpu(a)
    gpu = run_gpu(a)
    r = compare(cpu, gpu, a.rtol, a.atol)
    print(f"  final-field  max|dT|      : {r.max_abs:.6g} K\n"
          f"  final-field  max rel dT   : {r.max_rel:.6g}\n"
          f"  peak_T_K   CPU/GPU/|d|    : {cpu.peak_T:.10g} / {gpu.peak_T:.10g} / {r.d_peak_T:.3g}\n"
          f"  peak_W_um  CPU/GPU/|d|    : {cpu.peak_W:.10g} / {gpu.peak_W:.10g} / {r.d_peak_W:.3g}\n"
          f"  steadyWsub CPU/GPU/|d|    : {cpu.steady_W_sub:.10g} / {gpu.steady_W_sub:.10g} / "
          f"{r.d_steady:.3g}\n"
          f"  tolerance    rtol={a.rtol:g} atol={a.atol:g} K  -> {'PASS' if r.passed else 'FAIL'}",
          flush=True)
    return 0 if r.passed else 1


if __name__ == "__main__":
    sys.exit(main())
