"""Numerical model behind the least-squares precision tiers (CPU only, numpy
and LAPACK; no library call).  A 4096×512 matrix U·diag(s)·Vᵀ with s
geometric from 1 to 1/κ:

* FP64 least-squares solvers (LAPACK QR, gelsd, gelsy) agree to 1e-10 at
  κ = 1e6 but not at κ = 1e8, where their spread is about 1e-9: a 1e-10
  bar against an FP64 solution is below FP64's own accuracy there.
* A model of the mixed factorization (FP64 Householder panels of 128
  columns and their T; both trailing-update products, Z = Yᵀ·A and
  A −= (Y·Tᵀ)·Z, with FP32 operands and FP32 accumulation, added to the
  FP64 matrix) has a backward error of a few 1e-7: κ times it is below 1
  at κ = 1e6, where refinement converges, and far above 1 at κ = 1e8,
  where the stationary refinement cannot contract.  With FP64 products the
  same model is backward stable (a sanity check of the model itself).
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg import lapack

M, N, B = 4096, 512, 128


def graded(kappa, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((M, N)))
    V, _ = np.linalg.qr(rng.standard_normal((N, N)))
    A = (U * np.logspace(0, -np.log10(kappa), N)) @ V.T
    return A, A @ rng.standard_normal(N) + 1e-6 * rng.standard_normal(M)


def blocked_qr(A, fp32_update):
    """Returns the panels' (Y, T) and R of the model factorization."""
    F = np.array(A, order="F")
    panels = []
    for j in range(0, N, B):
        yr, tau, _, info = lapack.dgeqrf(np.asfortranarray(F[j:, j:j + B]))
        assert info == 0
        Y = np.tril(yr, -1) + np.eye(M - j, B)
        T = np.zeros((B, B))
        for i in range(B):             # forward compact-WY T (dlarft)
            T[i, i] = tau[i]
            T[:i, i] = -tau[i] * T[:i, :i] @ (Y[:, :i].T @ Y[:, i])
        F[j:, j:j + B] = yr
        C = F[j:, j + B:]
        if C.size:
            YT = Y @ T.T               # Qᵀ = I − Y·Tᵀ·Yᵀ
            if fp32_update:
                Z = Y.astype(np.float32).T @ C.astype(np.float32)
                C -= (YT.astype(np.float32) @ Z).astype(np.float64)
            else:
                C -= YT @ (Y.T @ C)
        panels.append((j, Y, T))
    return panels, np.triu(F[:N, :N])


def backward_error(A, panels, R):
    QR = np.vstack([R, np.zeros((M - N, N))])
    for j, Y, T in reversed(panels):   # apply Q = I − Y·T·Yᵀ panel by panel
        QR[j:] -= Y @ (T @ (Y.T @ QR[j:]))
    return np.linalg.norm(QR - A) / np.linalg.norm(A)


def spread(A, b):
    q, r = np.linalg.qr(A)
    xs = [sl.solve_triangular(r, q.T @ b)] + [sl.lstsq(A, b, lapack_driver=d)[0] for d in ("gelsd", "gelsy")]
    return max(np.abs(x - y).max() / np.abs(y).max() for x in xs for y in xs if x is not y)


@pytest.mark.parametrize("kappa,agree", [(1e6, True), (1e8, False)])
def test_fp64_solvers_spread(kappa, agree):
    s = spread(*graded(kappa, 7))
    assert (s < 1e-10) == agree, s


def test_model_is_backward_stable_in_fp64():
    A, _ = graded(1e8, 8)
    panels, R = blocked_qr(A, fp32_update=False)
    assert backward_error(A, panels, R) < 1e-13


@pytest.mark.parametrize("kappa,contracts", [(1e6, True), (1e8, False)])
def test_mixed_factorization_backward_error(kappa, contracts):
    A, _ = graded(kappa, 9)
    panels, R = blocked_qr(A, fp32_update=True)
    e = backward_error(A, panels, R)
    assert 1e-8 < e < 1e-6, e          # FP32-level, not FP64-level
    assert (kappa * e < 1) == contracts, kappa * e
