"""The 2D QR at its benchmark configuration on the GPU: QR_2D of a
32768×4096 iid N(0,1) matrix with b = 128 (test/QR/test_qr_2d.cxx:292-297
checks the same factorization by its residual).  Checked without a dense
m×m Q: randomized probes of ‖A − QR‖/(‖A‖_F·n·ε) and ‖I − QᵀQ‖ with Q
applied on the GPU from the panel factors, R against LAPACK's dgeqrf of
the same matrix up to row signs, and the first panel against the oracle's
hh_recon_qr of the first 32768×128 panel."""
import numpy as np
import pytest
import scipy.linalg.lapack as lp
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def Q():
    from candmc_amd import qr
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return qr


def test_qr_2d_32768x4096_b128(Q):
    m, n, b = 32768, 4096, 128
    print("QR_2D 32768x4096 b=128: generating", flush=True)
    A = O.gaussian(m, n, 20240607)
    dA = Q.to_gpu(A)
    dA2 = dA.clone()
    Q.QR_2D(dA, m, m, n, b)
    Tt = Q.geqrf2d(dA2, m, m, n, b)            # the same driver, keeping Tᵀ per panel
    torch.cuda.synchronize()
    assert torch.equal(dA, dA2)                # deterministic: the same bits both times
    print("QR_2D done; probing", flush=True)
    # probes: A·x − Q·[R·x; 0] and Qᵀ·Q·v − v, Q from the panel factors
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(n, 2, generator=g, dtype=torch.float64, device="cuda")
    Ad = Q.to_gpu(A).reshape(n, m).T           # m×n view of the column-major storage
    Ax = Ad @ x
    YR = dA.reshape(n, m).T
    R = torch.triu(YR[:n, :n])
    W = torch.zeros(m, 2, dtype=torch.float64, device="cuda")
    W[:n] = R @ x
    Wc = W.T.contiguous().reshape(-1)
    Q.apply_qt(dA, m, Tt, m, n, b, Wc, m, 2, trans=False)
    d = Ax - Wc.reshape(2, m).T
    normA = float(torch.linalg.norm(Ad))
    res = float(torch.linalg.norm(d, dim=0).max()) / (normA * float(torch.linalg.norm(x, dim=0).max())) / (n * EPS)
    v = torch.randn(m, 2, generator=g, dtype=torch.float64, device="cuda")
    vc = v.T.contiguous().reshape(-1)
    wc = vc.clone()
    Q.apply_qt(dA, m, Tt, m, n, b, wc, m, 2, trans=False)
    Q.apply_qt(dA, m, Tt, m, n, b, wc, m, 2, trans=True)
    orth = float(torch.linalg.norm(wc - vc)) / float(torch.linalg.norm(vc))
    print(f"resid/(n eps) {res:.3e}  orth {orth:.3e}", flush=True)
    assert res < 1e-3 and orth < 1e-14
    # R against LAPACK dgeqrf of the same matrix, up to row signs
    YRh = Q.from_gpu(dA, m, n)
    qr, _, _, info = lp.dgeqrf(np.asfortranarray(A.copy()))
    assert info == 0
    Rg, Rl = np.triu(YRh[:n]), np.triu(qr[:n])
    sg = np.sign(np.diag(Rg) / np.diag(Rl))
    err = np.abs(Rg - sg[:, None] * Rl).max() / np.abs(Rl).max()
    print(f"R vs dgeqrf: {err:.3e}", flush=True)
    assert err < 1e-12
    # the first panel is hh_recon_qr of A's first 128 columns
    YRo, _, _ = O.hh_recon_qr(A[:, :b].copy(order="F"), b, 1)
    assert np.abs(YRh[:, :b] - YRo).max() / np.abs(YRo).max() < 1e-12
