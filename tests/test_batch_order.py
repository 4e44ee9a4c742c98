"""Host-only check of the batched launches' dispatch order
(candmc_amd/csrc/batch.cpp) in tests/cxx/test_batch_order.cpp: every tree
node and construct_Q1 item of every problem gets one block, and every block
waits only on blocks with smaller ids.  Compiled with hipcc as host code
against the built library; no GPU call is made."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
LIB = os.path.join(ROOT, "candmc_amd", "libcandmc_qr.so")


@pytest.mark.skipif(not os.path.exists(HIPCC) or not os.path.exists(LIB),
                    reason="needs hipcc and the built library")
def test_batch_dispatch_order(tmp_path):
    exe = str(tmp_path / "batch_order")
    cmd = [HIPCC, "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "candmc_amd", "csrc"),
           os.path.join(ROOT, "tests", "cxx", "test_batch_order.cpp"),
           "-L" + os.path.dirname(LIB), "-lcandmc_qr", "-Wl,-rpath," + os.path.dirname(LIB), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "batch order ok" in run.stdout
