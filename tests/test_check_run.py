"""The --check counterpart (tests/check_run.py): its deviation measures and
pass rule against planted differences on the CPU, and on the GPU the
reference's check of the device-resident run for case 0 and for extended
physics, at the FP64 bound (1e-9 K) and, for FP32, at 0.5 K."""
import numpy as np
import pytest

from tests import check_run as cr


def _res(T, pk=2000.0, W=100.0, Ws=90.0):
    return cr.RunResult(pk, W, Ws, np.asarray(T, dtype=np.float64))


def test_defaults_are_the_reference_tolerances():
    a = cr.parse([])
    assert (a.rtol, a.atol, a.case, a.physics, a.tend) == (1e-6, 1e-3, "0", "base", 1.5e-3)


def test_compare_measures_and_pass_rule():
    cpu = _res([[300.0, 0.5], [2000.0, 1000.0]])
    # a 0.3 K error at the 1000 K cell, an error at a sub-kelvin cell counted
    # against the floor of 1 K
    gpu = _res([[300.0, 0.5 + 2e-4], [2000.0, 1000.3]], pk=2000.1, W=101.0, Ws=90.0)
    r = cr.compare(cpu, gpu, rtol=1e-6, atol=1e-3)
    assert r.max_abs == pytest.approx(0.3)
    assert r.max_rel == pytest.approx(0.3 / 1000.0)
    assert (r.d_peak_T, r.d_peak_W, r.d_steady) == (pytest.approx(0.1), 1.0, 0.0)
    assert not r.passed
    assert cr.compare(cpu, gpu, rtol=3e-4, atol=0.0).passed     # rel bound alone
    assert cr.compare(cpu, gpu, rtol=0.0, atol=0.3).passed      # abs bound alone
    assert cr.compare(cpu, cpu, rtol=0.0, atol=0.0).passed


def test_invalid_device_exits_3():
    assert cr.main(["--device=99", "--tend=1e-6"]) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["--tend=2e-4", "--check-rtol=0", "--check-atol=1e-9"],
    ["--case=3.1", "--physics=ext", "--kmult=2.5", "--tend=2e-4", "--check-rtol=0",
     "--check-atol=1e-9"],
    ["--tend=2e-4", "--precision=32", "--check-rtol=0", "--check-atol=0.5"],
])
def test_gpu_run_passes_check(args, capsys):
    print(f"\n[case] check_run {' '.join(args)}", flush=True)
    rc = cr.main(args)
    out = capsys.readouterr().out
    print(out)
    assert rc == 0 and "PASS" in out
