"""CPU check of the FP64 kernels' division by launch constants: the
reciprocal plus one remainder correction equals IEEE division on every
operand tried (tests/native/div_rn_check.c), and the uncorrected product,
planted as the mistake, does not."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reciprocal_division_is_correctly_rounded(tmp_path):
    exe = tmp_path / "div_rn_check"
    subprocess.run(["gcc", "-std=c99", "-O2", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "native", "div_rn_check.c"), "-o", str(exe), "-lm"],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    bad, tot = map(int, r.stdout.split())
    assert r.returncode == 0 and bad == 0 and tot > 3e7, r.stdout
    r = subprocess.run([str(exe), "naive"], capture_output=True, text=True, timeout=120)
    bad, tot = map(int, r.stdout.split())
    assert r.returncode == 1 and bad > 0.01 * tot, r.stdout
