"""GPU test of the bounded waits between workgroups of one launch (TSQR tree
nodes waiting for their children, construct_Q1 items waiting for their
parents): with the poll bound forced down from the host every wait gives
up at once, every kernel still finishes, and the call reports
CQR_ERR_DEVICE through the stream synchronisation instead of returning a
silently wrong factorization; the flag is consumed, so the next call runs
normally and is correct."""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    from candmc_amd import qr
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return qr


def test_forced_wait_timeout_is_reported_then_cleared(Q):
    from candmc_amd._lib import CQRError
    m, n, b = 16384, 256, 128
    A = O.gaussian(m, n, 77)
    ctx = Q.context()
    dA = Q.to_gpu(A)
    print("forced timeout: QR_2D", m, n, b, flush=True)
    try:
        ctx.set_wait_limit(1)
        with pytest.raises(CQRError) as e:
            Q.QR_2D(dA, m, m, n, b)
        assert e.value.status == 4          # CQR_ERR_DEVICE
    finally:
        ctx.set_wait_limit(0)
    dA = Q.to_gpu(A)
    Q.QR_2D(dA, m, m, n, b)                 # syncs and checks the flag again
    R = np.triu(Q.from_gpu(dA, n, n, m))
    Rl = sl.qr(A, mode="r")[0][:n, :n]
    s = np.sign(np.diag(R)) * np.sign(np.diag(Rl))
    assert np.abs(R - s[:, None] * Rl).max() / np.abs(Rl).max() < 1e-12
    with pytest.raises(CQRError):
        ctx.set_wait_limit(-1)


def test_forced_wait_timeout_in_the_one_launch_solve(Q):
    """The one-launch triangular solve's blocks wait for the blocks they
    depend on with the same bounded wait: forced down, the deep blocks give
    up, every workgroup finishes, the sync reports it, and the next solve is
    correct."""
    from candmc_amd._lib import CQRError
    n = 1030                                # nine 128-row blocks, the last ragged
    A = O.gaussian(n + 50, n, 9)
    dA = Q.to_gpu(A)
    Q.QR_2D(dA, n + 50, n + 50, n, 64)
    R = np.triu(Q.from_gpu(dA, n + 50, n)[:n, :n])
    B = O.gaussian(n, 2, 10)
    ctx = Q.context()
    print("forced timeout: trsm_r", n, flush=True)
    try:
        ctx.set_wait_limit(1)
        with pytest.raises(CQRError) as e:
            Q.trsm_r(dA, n + 50, n, Q.to_gpu(B), n, 2)
        assert e.value.status == 4          # CQR_ERR_DEVICE
    finally:
        ctx.set_wait_limit(0)
    dB = Q.to_gpu(B)
    Q.trsm_r(dA, n + 50, n, dB, n, 2)
    want = sl.solve_triangular(R, B, lower=False)
    assert np.abs(Q.from_gpu(dB, n, 2) - want).max() / np.abs(want).max() < 1e-11
