"""The reference-signature C++ layer (include/candmc/CANDMC_QR.h), exercised
by a C++ program written like the reference's own tests: drand48 inputs,
host arrays, and the column-relative residual bound of
test/QR/test_hh_recon.cxx:263 and test/QR/test_qr_2d.cxx:297."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "test_dropin")


def test_dropin_program():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.count("Test successful.") == 4
    assert "TEST FAILED" not in p.stdout
