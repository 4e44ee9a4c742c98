"""GPU parity of every hot-path entry point against the CPU oracle.

Runs through the C-ABI (candmc_amd.qr → libcandmc_qr.so).  Tolerances are
FP64 round-off bounds: the GPU and the oracle compute the same
factorization with different summation orders, so results agree to a few
ulps times the problem's growth (stated per test, relative to the largest
entry of the compared quantity).
"""
import os

import numpy as np
import pytest
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def Q():
    from candmc_amd import qr
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return qr


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def lapack_signs(YR, m, k):
    """The reference's tests flip the last R entry of a square factorization
    before comparing with LAPACK (test/QR/test_qr_2d.cxx:135-137)."""
    if m == k:
        YR = YR.copy(order="F")
        YR[m - 1, k - 1] *= -1.0
    return YR


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---------------------------------------------------------------- a1 leaf
@pytest.mark.parametrize("m,b", [(256, 128), (200, 128), (128, 128), (512, 64), (333, 64), (1024, 32),
                                 (1000, 17), (64, 16), (5, 5), (700, 100)])
def test_local_tsqr_equals_dgeqrf(Q, m, b):
    A = O.gaussian(m, b, 100 + m + b)
    dA = Q.to_gpu(A)
    dT = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    tau = Q.local_tsqr(dA, dT, m, b, m)
    qr, tl = O.local_tsqr(A)
    YR = Q.from_gpu(dA, m, b)
    assert rel(YR, qr) < 1e-13                                 # [V\R] of dgeqrf
    assert np.abs(tau.cpu().numpy() - tl).max() < 1e-14
    T = Q.from_gpu(dT, b, b)
    assert rel(T, O.dlarft(O.unit_lower(qr, b), tl)) < 1e-13   # compact-WY T


def test_local_tsqr_large_block_via_tsqr_hr(Q):
    m, b = 5000, 64                                            # beyond one CU's leaf
    A = O.gaussian(m, b, 7)
    dA = Q.to_gpu(A)
    tau = Q.local_tsqr(dA, None, m, b, m)
    qr, tl = O.local_tsqr(A)
    assert rel(Q.from_gpu(dA, m, b), qr) < 1e-12
    assert np.abs(tau.cpu().numpy() - tl).max() < 1e-13


def test_local_tsqr_zero_column(Q):
    A = O.gaussian(100, 8, 3)
    A[:, 3] = 0.0                                              # dlarfg xnorm = 0 → tau = 0
    dA = Q.to_gpu(A)
    tau = Q.local_tsqr(dA, None, 100, 8, 100)
    qr, tl = O.local_tsqr(A)
    assert tau.cpu().numpy()[3] == 0.0 and tl[3] == 0.0
    assert rel(Q.from_gpu(dA, 100, 8), qr) < 1e-13


def test_local_tsqr_tiny_scale_rescaling(Q):
    A = O.gaussian(80, 6, 5) * 1e-300                          # dlarfg safe-minimum path
    dA = Q.to_gpu(A)
    tau = Q.local_tsqr(dA, None, 80, 6, 80)
    qr, tl = O.local_tsqr(A)
    assert rel(Q.from_gpu(dA, 80, 6), qr) < 1e-12
    assert np.abs(tau.cpu().numpy() - tl).max() < 1e-13


# ---------------------------------------------------------------- a2 tree node
def test_tree_tsqr_golden(Q):
    g = gold("tpqrt_b20.npz")
    b = 20
    buf = np.vstack([g["R1"], g["R2"]])
    dA = Q.to_gpu(buf)
    dT = torch.zeros(16 * b, dtype=torch.float64, device="cuda")
    Q.tree_tsqr(dA, None, b, dT)
    out = Q.from_gpu(dA, 2 * b, b)
    assert rel(out[:b], g["R"]) < 1e-13
    assert rel(out[b:], g["V2"]) < 1e-13
    T = Q.from_gpu(dT, 16, b)
    # blocked T: column c holds rows s..c of its 16-wide block (s = 16·(c//16))
    keep = np.arange(16)[:, None] <= (np.arange(b)[None, :] % 16)
    assert rel(T * keep, g["T"] * keep) < 1e-13


@pytest.mark.parametrize("b", [128, 64, 33])
def test_tree_tsqr_random(Q, b):
    R1 = np.triu(O.gaussian(b, b, 11))
    R2 = np.triu(O.gaussian(b, b, 12))
    dA = Q.to_gpu(np.vstack([R1, R2]))
    Q.tree_tsqr(dA, None, b, None)
    R, V2, T = O.tree_tsqr(R1, R2, b)
    out = Q.from_gpu(dA, 2 * b, b)
    assert rel(out[:b], R) < 1e-13 and rel(out[b:], V2) < 1e-12


# ---------------------------------------------------------------- a3/a4 bitree + construct_Q1
@pytest.mark.parametrize("m,b,P", [(8192, 64, 4), (32768, 128, 0), (1000, 50, 3), (3000, 17, 0), (640, 64, 5),
                                   (2048, 40, 6), (1024, 24, 4), (4096, 120, 7)])
def test_bitree_tsqr_and_construct_Q1(Q, m, b, P):
    A = O.hh_recon_test_matrix(m, b)
    dA = Q.to_gpu(A)
    L, _, _ = Q.plan_tsqr(m, b, P)
    dR = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(L * b, dtype=torch.float64, device="cuda")
    td = torch.zeros(Q.tree_data_size(m, b, P), dtype=torch.float64, device="cuda")
    Q.bitree_tsqr(dA, m, dR, dtau, m, b, numPes=P, output_Y=1, tree_data=td)
    R = Q.from_gpu(dR, b, b)
    qr, _ = O.local_tsqr(A)
    Rl = np.triu(qr[:b])
    sg = np.sign(np.diag(R) / np.diag(Rl))
    assert rel(R, sg[:, None] * Rl) < 1e-12                   # R up to row signs
    dQ1 = torch.zeros(m * b, dtype=torch.float64, device="cuda")
    Q.construct_Q1(dA, m, dtau, dQ1, m, m, b, b, numPes=P, tree_data=td)
    Q1 = Q.from_gpu(dQ1, m, b)
    assert np.abs(Q1.T @ Q1 - np.eye(b)).max() < 1e-13       # ‖I − QᵀQ‖
    assert rel(Q1 @ R, A) < 1e-13                              # Q1·R = A


# ---------------------------------------------------------------- a5 signed LU
@pytest.mark.parametrize("b", [6, 64, 128])
def test_signed_nlu(Q, b):
    Qm, _ = np.linalg.qr(O.gaussian(3 * b, b, 21))
    A = np.asfortranarray(Qm[:b, :b].copy())
    R = np.triu(O.gaussian(b, b, 22))
    dA, dR = Q.to_gpu(A), Q.to_gpu(R)
    sg = torch.zeros(b, dtype=torch.int64, device="cuda")
    Q.recursive_NLU(dA, dR, b, b, b, sg)
    Ao, Ro, so = A.copy(order="F"), R.copy(order="F"), np.zeros(b, dtype=np.int64)
    O.recursive_NLU(Ao, Ro, so)
    assert np.array_equal(sg.cpu().numpy(), so)
    assert rel(Q.from_gpu(dA, b, b), Ao) < 1e-12
    assert np.array_equal(Q.from_gpu(dR, b, b), Ro)


# ---------------------------------------------------------------- a6/a7 hh_recon
@pytest.mark.parametrize("m,b,P,kind", [(136, 17, 4, "golden"), (8192, 64, 4, "drand"), (8192, 64, 4, "gauss"),
                                        (32768, 128, 0, "gauss"),
                                        (4096, 128, 0, "gauss"), (2000, 40, 3, "gauss"), (128, 128, 0, "gauss"),
                                        (16384, 16, 0, "gauss"), (777, 31, 2, "gauss")])
def test_hh_recon_qr(Q, m, b, P, kind):
    if kind == "golden":
        A = gold("hh_recon_136x17.npz")["A"]
    elif kind == "drand":
        A = O.hh_recon_test_matrix(m, b)
    else:
        A = O.gaussian(m, b, m * 7 + b)
    dA = Q.to_gpu(A)
    dW = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dT = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dI = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    Q.hh_recon_qr(dA, m, m, b, dW, numPes=P, T=dT, invT_T=dI)
    YR = Q.from_gpu(dA, m, b)
    YRo, Wo, _ = O.hh_recon_qr(A, b, P=max(P, 1) if m % b == 0 else 1)
    assert rel(YR, YRo) < 1e-12                                # [Y\R] (= dgeqrf's [V\R])
    assert rel(Q.from_gpu(dW, b, b), Wo) < 1e-12               # W = U·S
    Y = O.unit_lower(YRo, b)
    # HR reflectors have tau_i = 2/(1+‖y_i‖²) (for m == b the last one is the
    # sign flip tau = 2, which tau_recon would map to 0)
    tau = 2.0 / (1.0 + np.sum(np.tril(YRo, -1) ** 2, axis=0))
    assert rel(Q.from_gpu(dT, b, b), O.dlarft(Y, tau)) < 1e-12
    assert rel(Q.from_gpu(dI, b, b), O.compute_invT_from_W(Wo, Y[:b, :b])) < 1e-11
    assert O.column_relative_residual(A, lapack_signs(YR, m, b)) < 1e-9   # test_hh_recon.cxx:263


def test_hh_recon_same_for_any_leaf_count(Q):
    m, b = 4096, 64
    A = O.gaussian(m, b, 99)
    outs = []
    for P in (1, 8, 32):
        dA = Q.to_gpu(A)
        dW = torch.zeros(b * b, dtype=torch.float64, device="cuda")
        Q.hh_recon_qr(dA, m, m, b, dW, numPes=P)
        outs.append(Q.from_gpu(dA, m, b))
    assert rel(outs[1], outs[0]) < 1e-12 and rel(outs[2], outs[0]) < 1e-12


def test_compute_invT_from_W_and_Y(Q):
    m, b = 2976, 48                                            # 62 blocks of b rows
    A = O.gaussian(m, b, 31)
    YRo, Wo, _ = O.hh_recon_qr(A, b, 2)
    Y = O.unit_lower(YRo, b)
    d = Q.to_gpu(Y[:b, :b])
    Q.compute_invT_from_W(Q.to_gpu(Wo), b, d)
    want = O.compute_invT_from_W(Wo, Y[:b, :b])
    assert rel(Q.from_gpu(d, b, b), want) < 1e-12
    dT = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    Q.compute_invT_from_Y(Q.to_gpu(Y), m, dT, b, m, b)
    assert rel(Q.from_gpu(dT, b, b), O.compute_invT_from_Y(Y)) < 1e-12


# ---------------------------------------------------------------- a9/a10 updates
@pytest.mark.parametrize("mode", ["W", "Y", "T"])
def test_upd_A(Q, mode):
    m, b, k = 2496, 64, 300                                    # 39 blocks of b rows
    A = O.gaussian(m, b + k, 41)
    YRo, Wo, _ = O.hh_recon_qr(A[:, :b], b, 2)
    Y = O.unit_lower(YRo, b)
    invTT = O.compute_invT_from_W(Wo, Y[:b, :b])
    C = A[:, b:].copy(order="F")
    want = O.upd_A(Y, C, invTT)
    dC = Q.to_gpu(C)
    W = {"W": Q.to_gpu(Wo), "Y": None, "T": Q.to_gpu(invTT)}[mode]
    Q.upd_A(Q.to_gpu(Y), m, dC, m, m, k, b, W, W_is_T=(mode == "T"))
    assert rel(Q.from_gpu(dC, m, k), want) < 1e-12


def test_update_A_with_aggregation(Q):
    m, b, k = 1800, 32, 200
    A = O.gaussian(m, b + k, 43)
    YRo, Wo, _ = O.hh_recon_qr(A[:, :b], b, 1)
    Y = O.unit_lower(YRo, b)
    want = O.upd_A(Y, A[:, b:], O.compute_invT_from_W(Wo, Y[:b, :b]))
    dC = Q.to_gpu(A[:, b:])
    dAg = torch.zeros(m * b, dtype=torch.float64, device="cuda")
    Q.update_A(Q.to_gpu(YRo), m, dC, m, m, k, b, Q.to_gpu(Wo), aggreg_Y=dAg, lda_aY=m)
    assert rel(Q.from_gpu(dC, m, k), want) < 1e-12
    assert np.array_equal(Q.from_gpu(dAg, m, b), Y)


# ---------------------------------------------------------------- a11/a12 2D QR
@pytest.mark.parametrize("m,k,b", [(160, 96, 8), (160, 96, 32), (1024, 512, 128), (3000, 700, 64),
                                   (2048, 2048, 128), (999, 333, 50)])
def test_qr_2d(Q, m, k, b):
    A = gold("qr2d_gauss_160x96.npz")["A"] if m == 160 else O.gaussian(m, k, m + k + b)
    dA = Q.to_gpu(A)
    Q.QR_2D(dA, m, m, k, b)
    YR = Q.from_gpu(dA, m, k)
    want = O.QR_2D(A, b)
    assert rel(YR, want) < 1e-11
    assert O.column_relative_residual(A, lapack_signs(YR, m, k)) < 1e-9   # test_qr_2d.cxx:297
    if m > k:
        qr, _ = O.local_tsqr(A)
        assert rel(YR, qr) < 1e-11                             # = dgeqrf for m > k


@pytest.mark.parametrize("m,k,b", [(1500, 900, 64), (1024, 1024, 128)])
def test_qr_2d_lookahead_on_and_off(Q, m, k, b):
    """The two-stream lookahead order (qr_2d.cxx:378-635) and the plain
    right-looking order give the same factorization up to round-off."""
    A = O.gaussian(m, k, 7 * m + k)
    want = O.QR_2D(A, b)
    ctx = Q.context()
    out = []
    for la in (False, True):
        ctx.set_lookahead(la)
        try:
            dA = Q.to_gpu(A)
            Q.QR_2D(dA, m, m, k, b)
            out.append(Q.from_gpu(dA, m, k))
        finally:
            ctx.set_lookahead(True)
    for YR in out:
        assert rel(YR, want) < 1e-11
    assert O.column_relative_residual(A, lapack_signs(out[1], m, k)) < 1e-9
    assert rel(out[0], out[1]) < 1e-12


@pytest.mark.parametrize("m,n,k", [(4099, 700, 128), (32768, 520, 128), (777, 333, 64), (300, 129, 100),
                                   (8193, 4000, 128)])
def test_rank_b_update_kernel(Q, m, n, k):
    """upd_A with T⁻ᵀ = I (so Tᵀ = I): C −= Y·(Yᵀ·C), on full and ragged
    tiles of the rank-b update kernel, against an FP64 product on the host.
    (The kernel's overwrite form, Y = Q1·U⁻¹ and Y·Tᵀ, runs in every panel
    of the hh_recon_qr and QR_2D tests.)"""
    rng = np.random.default_rng(m + n + k)
    Y = np.asfortranarray(rng.standard_normal((m, k)) / np.sqrt(m))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    dC = Q.to_gpu(C)
    Q.upd_A(Q.to_gpu(Y), m, dC, m, m, n, k, Q.to_gpu(np.eye(k)), W_is_T=True)
    want = C - Y @ (Y.T @ C)
    assert rel(Q.from_gpu(dC, m, n), want) < 1e-12


def test_qr_2d_reference_default_input(Q):
    g = gold("qr2d_96x48.npz")                                 # singular: residual only
    dA = Q.to_gpu(g["A"])
    Q.QR_2D(dA, 96, 96, 48, 8)
    assert O.column_relative_residual(g["A"], Q.from_gpu(dA, 96, 48)) < 1e-9


def test_qr_2d_2d(Q):
    m, k = 2000, 600
    A = O.gaussian(m, k, 51)
    dA = Q.to_gpu(A)
    Q.QR_2D_2D(dA, m, m, k, 128, 32)
    want = O.QR_2D_2D(A, 128, 32)
    assert rel(Q.from_gpu(dA, m, k), want) < 1e-11


# ---------------------------------------------------------------- a14 helpers
def test_tau_recon_and_packing(Q):
    A = O.gaussian(300, 20, 61)
    qr, tau = O.local_tsqr(A)
    dt = torch.zeros(20, dtype=torch.float64, device="cuda")
    Q.tau_recon("L", 20, 300, 300, Q.to_gpu(qr), dt)
    assert np.abs(dt.cpu().numpy() - tau).max() < 1e-15
    R = np.triu(qr[:20, :20])
    dp = torch.zeros(O.psz_upr(20), dtype=torch.float64, device="cuda")
    Q.pack_upper(Q.to_gpu(R), dp, 20, 20)
    assert np.array_equal(dp.cpu().numpy(), O.pack_upper(R, 20))
    dR = torch.zeros(400, dtype=torch.float64, device="cuda")
    Q.unpack_upper(dp, dR, 20, 20)
    assert np.array_equal(Q.from_gpu(dR, 20, 20), R)
    Lm = np.tril(O.gaussian(20, 20, 62))
    dl = torch.zeros(O.psz_upr(20), dtype=torch.float64, device="cuda")
    Q.pack_lower(Q.to_gpu(Lm), dl, 20, 20, 1)
    assert np.array_equal(dl.cpu().numpy(), O.pack_lower(Lm, 20, 1))


# ---------------------------------------------------------------- error behaviour
def test_errors_are_status_codes(Q):
    from candmc_amd import CQRError
    d = torch.zeros(100, dtype=torch.float64, device="cuda")
    with pytest.raises(CQRError) as e:
        Q.hh_recon_qr(d, 5, 10, 20, d)                         # lda < m, m < b
    assert e.value.status == 1
    with pytest.raises(CQRError) as e:
        Q.QR_2D(torch.zeros(200 * 200, dtype=torch.float64, device="cuda"), 200, 200, 200, 256)
    assert e.value.status == 2
    ctx = Q.context()
    ctx.set_check_finite(True)
    try:
        A = O.gaussian(300, 16, 1)
        A[7, 3] = np.nan
        with pytest.raises(CQRError) as e:
            Q.hh_recon_qr(Q.to_gpu(A), 300, 300, 16, torch.zeros(256, dtype=torch.float64, device="cuda"))
        assert e.value.status == 5
    finally:
        ctx.set_check_finite(False)
