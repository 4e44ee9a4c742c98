"""The allocation-free panel factorization cqr_hh_recon_qr_noalloc (no
counterpart in the reference; for callers that cannot allocate) against
cqr_hh_recon_qr_ex through the C-ABI: the same kernels on a fixed
per-GPU scratch array give bit-identical [Y\\R], W, T and T⁻ᵀ on two calls
in a row, the second smaller; a panel too large for the scratch is refused
with CQR_ERR_UNSUPPORTED and left untouched; the context's workspace arena
does not grow."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu


def _pd(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(lib, h, A, m, b, P, noalloc):
    dA = torch.from_numpy(np.ascontiguousarray(A.reshape(-1, order="F"))).to("cuda")
    W, T, I = (torch.zeros(b * b, dtype=torch.float64, device="cuda") for _ in range(3))
    fn = lib.cqr_hh_recon_qr_noalloc if noalloc else lib.cqr_hh_recon_qr_ex
    st = fn(h, _pd(dA), m, m, b, _pd(W), _pd(T), _pd(I), P)
    assert st == 0, lib.cqr_last_error(h).decode()
    assert lib.cqr_sync(h) == 0
    return [x.cpu().numpy() for x in (dA, W, T, I)]


def test_noalloc_matches_allocating_entry_point():
    from candmc_amd import _lib
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.cqr_create(ctypes.byref(h), 0) == 0
    try:
        assert lib.cqr_noalloc_scratch_bytes() == 96 << 20
        for m, b, P in ((16384, 128, 0), (3000, 64, 3)):     # two calls in a row, the second smaller
            print(f"noalloc m={m} b={b} P={P}", flush=True)
            A = O.gaussian(m, b, 700 + m)
            before = lib.cqr_workspace_bytes(h)
            got = _run(lib, h, A, m, b, P, True)
            assert lib.cqr_workspace_bytes(h) == before         # nothing allocated
            want = _run(lib, h, A, m, b, P, False)
            for g, w in zip(got, want):
                assert np.array_equal(g, w)
            YRo, Wo, _ = O.hh_recon_qr(A, b, max(P, 1) if m % b == 0 else 1)
            YR = got[0].reshape(b, m).T
            assert np.abs(YR - YRo).max() / np.abs(YRo).max() < 1e-12
        # too large for the scratch: refused, A untouched
        m, b = 24576, 128
        A = torch.ones(m * b, dtype=torch.float64, device="cuda")
        W = torch.zeros(b * b, dtype=torch.float64, device="cuda")
        st = lib.cqr_hh_recon_qr_noalloc(h, _pd(A), m, m, b, _pd(W), None, None, 0)
        assert st == 2
        assert lib.cqr_sync(h) == 0
        assert bool((A == 1.0).all())
    finally:
        lib.cqr_destroy(h)
