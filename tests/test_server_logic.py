"""Host-only check of the request server's queue and collection logic
(candmc_amd/csrc/server.cpp) with its GPU calls replaced by stand-ins in
tests/cxx/test_server_logic.cpp; the serving loop has no counterpart in
the reference.  Built with g++ (no GPU code), run under a time limit so a
wait that never returns fails the test instead of hanging it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(ROCM_INC),
                    reason="needs g++ and the HIP headers")
def test_server_queue_logic(tmp_path):
    exe = str(tmp_path / "server_logic")
    cmd = ["g++", "-O1", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INC,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "candmc_amd", "csrc"),
           os.path.join(ROOT, "candmc_amd", "csrc", "server.cpp"),
           os.path.join(ROOT, "tests", "cxx", "test_server_logic.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "server logic ok" in run.stdout
