"""The product path has no CPU fallback: without the built HIP library, a
GPU or a GPU tensor every call raises, and nothing in the package imports
the oracle."""
import os
import re

import pytest
import torch

from candmc_amd import _lib, qr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libcandmc_qr.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_host_tensors_are_rejected():
    with pytest.raises(TypeError, match="GPU tensor"):
        qr._p(torch.zeros(4, dtype=torch.float64))
    with pytest.raises(TypeError, match="float64"):
        qr._pd(torch.zeros(4, dtype=torch.float32))
    with pytest.raises(TypeError):
        qr._p([0.0, 1.0])


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the behaviour without a GPU")
def test_calls_raise_without_a_gpu():
    """The library's context needs a HIP device: creating it fails with
    CQR_ERR_DEVICE before any host tensor is looked at."""
    A = torch.zeros(64 * 16, dtype=torch.float64)
    for call in (lambda: qr.QR_2D(A, 64, 64, 16, 16), lambda: qr.QR_tree_2D(A, 64, 64, 16, 16)):
        with pytest.raises(_lib.CQRError) as e:
            call()
        assert e.value.status == 4
    assert not A.any()                   # nothing was computed on the host


def test_package_never_imports_the_oracle():
    pat = re.compile(r"^\s*(import\s+oracle|from\s+oracle\b)|candmc_oracle", re.M)
    for d, _, files in os.walk(os.path.join(ROOT, "candmc_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(d, f), encoding="utf-8").read()
                assert not pat.search(src), os.path.join(d, f)


def test_library_links_no_vendor_blas():
    """Every product of the path runs on the library's own MFMA kernels:
    the shared library depends on no BLAS (rocBLAS, hipBLAS, hipBLASLt)."""
    import subprocess
    from candmc_amd import _lib
    out = subprocess.run(["readelf", "-d", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    needed = [l for l in out.splitlines() if "NEEDED" in l]
    assert needed
    assert not any(("blas" in l.lower()) for l in needed), needed
