"""GPU parity of the least-squares path: panel factors, Qᵀ/Q application,
triangular solves and the refined least-squares solve (FP64 and with the
FP32 trailing update), against the oracle and scipy's LAPACK lstsq.

Tolerances: FP64 results agree to round-off (1e-12 relative); refined
solutions agree with the FP64 reference to 1e-10 relative, or to a small
multiple of κ(A)·ε where that is larger (no FP64 solver does better)."""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def Q():
    from candmc_amd import qr
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return qr


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.mark.parametrize("m,k,b", [(1000, 300, 128), (700, 200, 64), (512, 512, 128)])
def test_geqrf2d_panel_factors(Q, m, k, b):
    """[Y\\R] equals QR_2D's; each Tt block is Tᵀ = dlarft(Y_p, τ_p)ᵀ."""
    A = O.gaussian(m, k, m + k)
    dA = Q.to_gpu(A)
    Tt = Q.geqrf2d(dA, m, m, k, b)
    YR = Q.from_gpu(dA, m, k)
    assert rel(YR, O.QR_2D(A, b)) < 1e-11
    tt = Tt.cpu().numpy()
    npan = (k + b - 1) // b
    for p in range(npan):
        j, bj = p * b, min(b, k - p * b)
        P = YR[j:, j:j + bj]
        T = O.dlarft(O.unit_lower(P, bj), O.tau_recon("L", P, bj, m - j))
        got = np.tril(tt[p * b * b: p * b * b + bj * bj].reshape(bj, bj, order="F"))
        # a square matrix's last reflector is the sign quirk of
        # test_qr_2d.cxx:135-137 (no subdiagonal part): compare the rest
        c = bj - 1 if (m == k and p == npan - 1) else bj
        assert rel(got[:c, :c], T.T[:c, :c]) < 1e-11


@pytest.mark.parametrize("nrhs", [1, 5, 16])
def test_apply_qt_and_q(Q, nrhs):
    m, k, b = 900, 260, 64
    A = O.gaussian(m, k, 3)
    B = O.gaussian(m, nrhs, 4)
    dA = Q.to_gpu(A)
    Tt = Q.geqrf2d(dA, m, m, k, b)
    YR = Q.from_gpu(dA, m, k)
    dB = Q.to_gpu(B)
    Q.apply_qt(dA, m, Tt, m, k, b, dB, m, nrhs, trans=True)
    QtB = Q.from_gpu(dB, m, nrhs)
    assert rel(QtB, O.apply_QT(YR, B)) < 1e-12
    Q.apply_qt(dA, m, Tt, m, k, b, dB, m, nrhs, trans=False)    # Q·(Qᵀ·B) = B
    assert rel(Q.from_gpu(dB, m, nrhs), B) < 1e-13


# n ≥ 384 (four 128-row blocks) takes the one-launch solve, shorter ones the
# per-block launches; 1030 leaves a 6-row last block
# panels of at least 16384 rows with at most four columns sum their slab
# partials and apply T in one workgroup
@pytest.mark.parametrize("m,k,b,nrhs", [(20000, 260, 64, 1), (17000, 300, 128, 4)])
def test_apply_qt_and_q_tall(Q, m, k, b, nrhs):
    A = O.gaussian(m, k, 5)
    B = O.gaussian(m, nrhs, 6)
    dA = Q.to_gpu(A)
    Tt = Q.geqrf2d(dA, m, m, k, b)
    YR = Q.from_gpu(dA, m, k)
    dB = Q.to_gpu(B)
    Q.apply_qt(dA, m, Tt, m, k, b, dB, m, nrhs, trans=True)
    assert rel(Q.from_gpu(dB, m, nrhs), O.apply_QT(YR, B)) < 1e-12
    Q.apply_qt(dA, m, Tt, m, k, b, dB, m, nrhs, trans=False)    # Q·(Qᵀ·B) = B
    assert rel(Q.from_gpu(dB, m, nrhs), B) < 1e-13


@pytest.mark.parametrize("n,trans,nrhs", [(300, False, 3), (300, True, 3), (128, False, 3), (517, True, 3),
                                          (517, False, 1), (640, True, 16), (1030, False, 16), (1030, True, 1)])
def test_trsm_r(Q, n, trans, nrhs):
    A = O.gaussian(n + 50, n, 9)
    dA = Q.to_gpu(A)
    Q.QR_2D(dA, n + 50, n + 50, n, 64)
    YR = Q.from_gpu(dA, n + 50, n)
    R = np.triu(YR[:n, :n])
    B = O.gaussian(n, nrhs, 10)
    dB = Q.to_gpu(B)
    Q.trsm_r(dA, n + 50, n, dB, n, nrhs, trans=trans)
    want = sl.solve_triangular(R, B, trans=1 if trans else 0, lower=False)
    assert rel(Q.from_gpu(dB, n, nrhs), want) < 1e-11


@pytest.mark.parametrize("m,n,nrhs,b", [(2000, 300, 1, 128), (1500, 256, 4, 64), (3000, 400, 16, 128)])
def test_lstsq_fp64(Q, m, n, nrhs, b):
    A = O.gaussian(m, n, 11 + n)
    B = O.gaussian(m, nrhs, 12 + n)
    X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=b, max_iter=3, tol=1e-12)
    x = Q.from_gpu(X, n, nrhs)
    want = sl.lstsq(A, B)[0]
    assert rel(x, want) < 1e-12 and upd <= 1e-12
    assert np.abs(A.T @ (B - A @ x)).max() < 1e-10 * np.abs(A).max() * np.abs(B).max() * m
    assert 1 <= iters <= 3


# from 8192×1024 the residual products A·x split their columns over eight
# workgroup rows; 16500 rows also reach the one-workgroup T product of the Q
# application, 1100 columns a ragged last block of the one-launch solve
@pytest.mark.parametrize("m,n,nrhs", [(8192, 1024, 1), (16500, 1100, 5)])
def test_lstsq_fp64_tall_wide(Q, m, n, nrhs):
    A = O.gaussian(m, n, 21)
    B = O.gaussian(m, nrhs, 22)
    X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, max_iter=3, tol=1e-12)
    x = Q.from_gpu(X, n, nrhs)
    want = sl.lstsq(A, B)[0]
    assert rel(x, want) < 1e-12 and upd <= 1e-12
    assert 1 <= iters <= 3


def _conditioned_problem(m, n, nrhs, kappa, seed):
    A = O.conditioned_matrix(m, n, kappa, seed)
    rng = np.random.default_rng(seed + 1)
    B = np.asfortranarray(A @ rng.standard_normal((n, nrhs)) + 1e-6 * rng.standard_normal((m, nrhs)))
    return A, B, sl.lstsq(A, B)[0]


def test_lstsq_fp32_update_refined_to_fp64(Q):
    """κ = 1e2: the FP32 trailing update + FP64-residual refinement reaches
    the FP64 least-squares solution without refactoring."""
    m, n, nrhs = 4096, 256, 2
    A, B, want = _conditioned_problem(m, n, nrhs, 1e2, 21)
    X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_F32,
                            max_iter=20, tol=1e-12)
    assert Q.context().last_prec() == Q.PREC_F32
    assert rel(Q.from_gpu(X, n, nrhs), want) < 1e-10, (iters, upd)
    assert iters >= 2 and upd <= 1e-12    # the FP32 factorization alone is not FP64-accurate
    X0, _, _ = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_F32, max_iter=0)
    assert rel(Q.from_gpu(X0, n, nrhs), want) > 1e-12


def test_lstsq_fp32_at_kappa_1e6_with_and_without_escalation(Q):
    """κ = 1e6 (κ·u₃₂ ≈ 0.06 < 1): the FP32 factors refine to the FP64
    solution, with escalation on (it may refactor in FP64 when R's diagonal
    range flags the factors) and off (always the FP32 factors).  Either way
    the answer is within max(1e-10, 20κε) and the status is CQR_OK."""
    m, n, nrhs, kappa = 4096, 256, 2, 1e6
    A, B, want = _conditioned_problem(m, n, nrhs, kappa, 23)
    ctx = Q.context()
    X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_F32,
                            max_iter=30, tol=1e-10)
    assert ctx.last_prec() in (Q.PREC_F32, Q.PREC_F64) and upd <= 1e-10
    assert rel(Q.from_gpu(X, n, nrhs), want) < max(1e-10, 20 * kappa * EPS), (iters, upd)
    ctx.set_lstsq_escalate(False)
    try:
        X, it2, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_F32,
                              max_iter=30, tol=1e-10)
        assert ctx.last_prec() == Q.PREC_F32 and upd <= 1e-10
    finally:
        ctx.set_lstsq_escalate(True)
    assert rel(Q.from_gpu(X, n, nrhs), want) < max(1e-10, 20 * kappa * EPS), (it2, upd)
    assert it2 >= 2


@pytest.mark.parametrize("prec,kappa", [(1, 1e8), (2, 1e8), (2, 1e6)])
def test_lstsq_mixed_never_fails_silently(Q, prec, kappa):
    """The mixed configurations whose refinement cannot reach 1e-10 (κ·u of
    the factorization ≳ 1): with escalation on, the solve is refactored in
    FP64 and either meets 1e-10 against scipy's FP64 solution or returns
    CQR_ERR_NOCONVERGE with FP64-quality X; with escalation off it either
    meets 1e-10 or returns CQR_ERR_NOCONVERGE.  CQR_OK with an error above
    1e-10 fails the test."""
    m, n, nrhs = 4096, 256, 1
    A, B, want = _conditioned_problem(m, n, nrhs, kappa, 41 + prec)
    ctx = Q.context()
    dA, dB = Q.to_gpu(A), Q.to_gpu(B)
    X, iters, upd, st = Q.lstsq(dA, m, m, n, dB, m, nrhs, b=128, prec=prec, max_iter=30, tol=1e-10,
                                raise_noconverge=False)
    err = rel(Q.from_gpu(X, n, nrhs), want)
    assert ctx.last_prec() == Q.PREC_F64
    assert (st == 0 and err <= 1e-10 and upd <= 1e-10) or (st == 6 and upd > 1e-10), (st, err, upd)
    assert err < 100 * kappa * EPS       # FP64-quality either way
    ctx.set_lstsq_escalate(False)
    try:
        X, iters, upd, st = Q.lstsq(dA, m, m, n, dB, m, nrhs, b=128, prec=prec, max_iter=30, tol=1e-10,
                                    raise_noconverge=False)
        assert ctx.last_prec() == prec
    finally:
        ctx.set_lstsq_escalate(True)
    err = rel(Q.from_gpu(X, n, nrhs), want)
    assert (st == 0 and err <= 1e-10) or st == 6, (st, err, upd)


def test_lstsq_errors(Q):
    from candmc_amd._lib import CQRError
    A = Q.to_gpu(O.gaussian(100, 10, 1))
    B = Q.to_gpu(O.gaussian(100, 17, 2))
    with pytest.raises(CQRError) as e:
        Q.lstsq(A, 100, 100, 10, B, 100, 17)
    assert e.value.status == 2


def test_lstsq_bf16_update_refined_to_fp64(Q):
    """BF16 trailing update (u = 2⁻⁸) converges under refinement when
    κ·u < 1: κ = 1e2 reaches the FP64 solution, in more steps than FP32
    (escalation off: with it on, κ̂·u > 1e-2 refactors in FP64 at once)."""
    m, n, nrhs = 4096, 256, 1
    A, B, want = _conditioned_problem(m, n, nrhs, 1e2, 31)
    ctx = Q.context()
    ctx.set_lstsq_escalate(False)
    try:
        X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_BF16,
                                max_iter=80, tol=1e-12)
        assert ctx.last_prec() == Q.PREC_BF16
        _, it32, _ = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_F32,
                             max_iter=80, tol=1e-12)
    finally:
        ctx.set_lstsq_escalate(True)
    assert rel(Q.from_gpu(X, n, nrhs), want) < 1e-10, (iters, upd)
    assert it32 < iters < 80
    X, iters, upd = Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(B), m, nrhs, b=128, prec=Q.PREC_BF16,
                            max_iter=80, tol=1e-12)
    assert ctx.last_prec() == Q.PREC_F64
    assert rel(Q.from_gpu(X, n, nrhs), want) < 1e-10


def test_lstsq_nan_in_any_column_is_reported(Q):
    """A NaN in the second right-hand side (not only the first) ends the
    solve with CQR_ERR_NONFINITE instead of CQR_OK."""
    from candmc_amd._lib import CQRError
    m, n = 600, 40
    A = O.gaussian(m, n, 51)
    B = O.gaussian(m, 2, 52)
    B[7, 1] = np.nan
    with pytest.raises(CQRError) as e:
        Q.lstsq(Q.to_gpu(A), m, m, n, Q.to_gpu(np.asfortranarray(B)), m, 2, b=128, max_iter=4, tol=1e-12)
    assert e.value.status == 5
