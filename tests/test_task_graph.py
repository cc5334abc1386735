"""CPU check of the Hedgehog-free task graph behind the step-task interface."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_graph_delivers_every_result_and_propagates_errors(tmp_path):
    exe = tmp_path / "graph_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "graph_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr


def test_hedgehog_adapter_compiles_against_the_task_interface(tmp_path):
    """include/lpbf/hh_step_task.hpp against a stand-in of the Hedgehog task
    interface the reference's StepTileTask uses; construction and copy()
    (no GPU work: a task creates its device state on first use)."""
    exe = tmp_path / "hh_adapter_check"
    lib = os.path.join(ROOT, "hedgehog_lpbf_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "tests", "native", "hh_standin"),
                    os.path.join(ROOT, "tests", "native", "hh_adapter_check.cpp"), "-o", str(exe),
                    "-L", lib, "-lhedgehog_lpbf_amd", "-Wl,-rpath," + lib],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
