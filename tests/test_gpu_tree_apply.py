"""GPU parity of the implicit TSQR Q and its callers against the CPU oracle:
apply_tsqr_QT (alg/QR/tsqr/apply_tsqr_QT.cxx:31-165) and its Q direction,
construct_Q1 for k ≠ b and with ID (alg/QR/tsqr/construct_tsqr_Q.cxx:34-242),
QR_tree_2D (alg/QR/qr_2d/qr_tree_2d.cxx:30-124) with its [A, A] invariant
(test/QR/test_qr_tree_2d.cxx:83-143), and the export of tree_data in the
reference's layout (alg/QR/tsqr/bitree_tsqr.cxx:213-217).

The oracle factors the same row blocks as the GPU (the leaf split that
cqr_plan_tsqr reports), so both compute the same Q; tolerances are FP64
round-off bounds relative to the largest entry of the compared quantity.
"""
import numpy as np
import pytest
import torch

from oracle import candmc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Q():
    from candmc_amd import qr
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return qr


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def offsets(m, L):
    return [(l * m) // L for l in range(L + 1)]


def factor(Q, A, b, P):
    """GPU bitree TSQR of A (m×b) and the oracle's on the same leaves."""
    m = A.shape[0]
    L, _, _ = Q.plan_tsqr(m, b, P)
    dA = Q.to_gpu(A)
    dR = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(L * b, dtype=torch.float64, device="cuda")
    td = torch.zeros(Q.tree_data_size(m, b, P), dtype=torch.float64, device="cuda")
    Q.bitree_tsqr(dA, m, dR, dtau, m, b, numPes=P, output_Y=1, tree_data=td)
    o = offsets(m, L)
    Ys, R, taus, otd = O.bitree_tsqr([A[o[r]:o[r + 1]].copy(order="F") for r in range(L)], b)
    return dict(dY=dA, dR=dR, dtau=dtau, td=td, L=L, o=o, Ys=Ys, R=R, taus=taus, otd=otd)


def blocks(X, o):
    return [X[o[r]:o[r + 1]].copy(order="F") for r in range(len(o) - 1)]


@pytest.mark.parametrize("m,b,P,k", [(1000, 50, 3, 37), (2048, 64, 0, 200), (4096, 128, 0, 300),
                                     (640, 64, 5, 64), (3000, 17, 0, 5), (8192, 32, 0, 1)])
def test_apply_tsqr_QT_and_Q(Q, m, b, P, k):
    print(f"apply_tsqr m={m} b={b} P={P} k={k}", flush=True)
    A = O.gaussian(m, b, 500 + m + b)
    B = O.gaussian(m, k, 600 + k)
    f = factor(Q, A, b, P)
    dB = Q.to_gpu(B)
    Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dB, m, m, b, k, numPes=P, tree_data=f["td"])
    got = Q.from_gpu(dB, m, k)
    want = np.vstack(O.apply_tsqr_QT(f["Ys"], f["taus"], f["otd"], blocks(B, f["o"]), b))
    assert rel(got, want) < 1e-12
    # Q·(Qᵀ·B) = B
    Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dB, m, m, b, k, numPes=P, tree_data=f["td"], trans=False)
    assert rel(Q.from_gpu(dB, m, k), B) < 1e-12
    # Qᵀ·A = [R; 0], R at the top rows of leaf 0
    dA0 = Q.to_gpu(A)
    Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dA0, m, m, b, b, numPes=P, tree_data=f["td"])
    qta = Q.from_gpu(dA0, m, b)
    R = Q.from_gpu(f["dR"], b, b)
    z = np.zeros((m, b))
    z[:b] = R
    assert rel(qta, z) < 1e-12


@pytest.mark.parametrize("m,b,P", [(1000, 50, 3), (2048, 64, 0), (640, 64, 5), (300, 12, 1)])
def test_construct_Q1_any_k_and_ID(Q, m, b, P):
    print(f"construct_Q1 m={m} b={b} P={P}", flush=True)
    A = O.gaussian(m, b, 700 + m)
    f = factor(Q, A, b, P)
    for k in (b + 7, 2 * b, max(1, b // 2)):
        dQ1 = torch.zeros(m * k, dtype=torch.float64, device="cuda")
        Q.construct_Q1(f["dY"], m, f["dtau"], dQ1, m, m, b, k, numPes=P, tree_data=f["td"])
        want = np.vstack(O.construct_Q1_full(f["Ys"], f["taus"], f["otd"], b, k))
        got = Q.from_gpu(dQ1, m, k)
        assert rel(got, want) < 1e-12
        assert np.abs(got.T @ got - np.eye(k)).max() < 1e-12
    ID = O.gaussian(b, b, 9)
    dQ1 = torch.zeros(m * b, dtype=torch.float64, device="cuda")
    Q.construct_Q1(f["dY"], m, f["dtau"], dQ1, m, m, b, b, numPes=P, tree_data=f["td"], ID=Q.to_gpu(ID))
    want = np.vstack(O.construct_Q1_full(f["Ys"], f["taus"], f["otd"], b, b, ID=ID))
    assert rel(Q.from_gpu(dQ1, m, b), want) < 1e-12
    # the general path with ID = I equals the fused k = b kernel
    dI = torch.zeros(m * b, dtype=torch.float64, device="cuda")
    Q.construct_Q1(f["dY"], m, f["dtau"], dI, m, m, b, b, numPes=P, tree_data=f["td"], ID=Q.to_gpu(np.eye(b)))
    dF = torch.zeros(m * b, dtype=torch.float64, device="cuda")
    Q.construct_Q1(f["dY"], m, f["dtau"], dF, m, m, b, b, numPes=P, tree_data=f["td"])
    assert rel(Q.from_gpu(dI, m, b), Q.from_gpu(dF, m, b)) < 1e-13


def aa_defect(out, k):
    """test_qr_tree_2d.cxx:83-143: strictly-lower right half squared, upper
    part min(absolute, relative) difference to the left half."""
    m = out.shape[0]
    Lh, Rh = out[:, :k], out[:, k:2 * k]
    low = np.arange(m)[:, None] > np.arange(k)[None, :]
    d = Lh - Rh
    up = np.minimum(d * d, (d / np.where(Lh == 0, 1.0, Lh)) ** 2)
    return float(np.sqrt(np.sum(np.where(low, Rh * Rh, up))))


@pytest.mark.parametrize("m,k,b,stop", [(600, 300, 64, -1), (256, 256, 128, -1), (2000, 500, 128, 256),
                                        (900, 200, 50, -1)])
def test_qr_tree_2d_one_rank(Q, m, k, b, stop):
    print(f"qr_tree_2d one rank m={m} k={k} b={b} stop={stop}", flush=True)
    A = O.gaussian(m, k, 800 + m + k)
    dA = Q.to_gpu(A)
    Q.QR_tree_2D(dA, m, m, k, b, numPes=1, stop_at=stop)
    want = O.QR_tree_2D(A, b, stop_at=stop)
    assert rel(Q.from_gpu(dA, m, k), want) < 1e-11


@pytest.mark.parametrize("m,k,b,P", [(4096, 256, 64, 0), (2048, 192, 128, 0), (1200, 96, 32, 5)])
def test_qr_tree_2d_leaves_AA(Q, m, k, b, P):
    """The reference test's input and bar (test_qr_tree_2d.cxx:64-143).  Its
    drand48 lattice makes some leaf blocks numerically rank deficient, so
    reflector signs at round-off-sized pivots are arbitrary: only the
    invariant is compared on it."""
    print(f"qr_tree_2d leaves [A,A] m={m} k={k} b={b} P={P}", flush=True)
    A = O.qr_2d_test_matrix(m, k)
    dA = Q.to_gpu(np.hstack([A, A]))
    Q.QR_tree_2D(dA, m, m, 2 * k, b, numPes=P, stop_at=min(m, k))
    got = Q.from_gpu(dA, m, 2 * k)
    assert aa_defect(got, k) < 1e-6                  # the reference test's bar
    assert aa_defect(got, k) / np.abs(got[:k, :k]).max() < 1e-12


@pytest.mark.parametrize("m,k,b,P", [(4096, 256, 64, 0), (2048, 320, 128, 0), (1200, 96, 32, 5)])
def test_qr_tree_2d_leaves_vs_oracle(Q, m, k, b, P):
    print(f"qr_tree_2d leaves m={m} k={k} b={b} P={P}", flush=True)
    A = O.gaussian(m, k, 1000 + m + k)
    dA = Q.to_gpu(A)
    stop = min(m, k) - b // 2
    Q.QR_tree_2D(dA, m, m, k, b, numPes=P, stop_at=stop)
    got = Q.from_gpu(dA, m, k)

    def split(mm, bw):
        L, _, _ = Q.plan_tsqr(mm, bw, P)
        return np.diff(offsets(mm, L))

    want = O.QR_tree_2D(A, b, split=split, stop_at=stop)
    assert rel(got, want) < 1e-11


@pytest.mark.parametrize("m,b,P", [(2048, 64, 0), (1000, 20, 6), (4096, 128, 0)])
def test_tree_data_export_layout(Q, m, b, P):
    print(f"tree_data export m={m} b={b} P={P}", flush=True)
    A = O.gaussian(m, b, 900 + m)
    f = factor(Q, A, b, P)
    nb = min(b, 16)
    keep = np.arange(nb)[:, None] <= (np.arange(b)[None, :] % nb)
    for r in range(f["L"]):
        got = Q.tree_data_export(f["td"], m, b, P, r).cpu().numpy()
        want = O.tree_data_packed(f["otd"][r], b)
        assert got.size == want.size
        step = nb * b + b * (b + 1) // 2
        for lv in range(len(f["otd"][r])):
            g, w = got[lv * step:(lv + 1) * step], want[lv * step:(lv + 1) * step]
            Tg, Tw = g[:nb * b].reshape(b, nb).T, w[:nb * b].reshape(b, nb).T
            assert rel(Tg * keep, Tw * keep) < 1e-11
            assert rel(g[nb * b:], w[nb * b:]) < 1e-11


def test_construct_Q1_full_square_Q(Q):
    """k = m: the whole orthogonal Q of the TSQR (QᵀQ = I, Qᵀ·A = [R; 0])."""
    m, b, P = 600, 40, 4
    A = O.gaussian(m, b, 1234)
    f = factor(Q, A, b, P)
    dQ = torch.zeros(m * m, dtype=torch.float64, device="cuda")
    Q.construct_Q1(f["dY"], m, f["dtau"], dQ, m, m, b, m, numPes=P, tree_data=f["td"])
    Qm = Q.from_gpu(dQ, m, m)
    assert np.abs(Qm.T @ Qm - np.eye(m)).max() < 1e-13
    R = Q.from_gpu(f["dR"], b, b)
    z = np.zeros((m, b))
    z[:b] = R
    assert rel(Qm.T @ A, z) < 1e-13
    assert rel(Qm, np.vstack(O.construct_Q1_full(f["Ys"], f["taus"], f["otd"], b, m))) < 1e-12


def test_apply_tsqr_strided_and_empty(Q):
    """B with a leading dimension larger than m and k = 0 (a no-op)."""
    m, b, P, k, ldb = 1000, 50, 3, 9, 1031
    A = O.gaussian(m, b, 77)
    B = O.gaussian(m, k, 78)
    f = factor(Q, A, b, P)
    Bp = np.zeros((ldb, k))
    Bp[:m] = B
    dB = Q.to_gpu(Bp)
    Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dB, ldb, m, b, k, numPes=P, tree_data=f["td"])
    got = Q.from_gpu(dB, ldb, k)
    want = np.vstack(O.apply_tsqr_QT(f["Ys"], f["taus"], f["otd"], blocks(B, f["o"]), b))
    assert rel(got[:m], want) < 1e-12
    assert np.array_equal(got[m:], Bp[m:])                 # rows beyond m untouched
    Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dB, ldb, m, b, 0, numPes=P, tree_data=f["td"])
    assert np.array_equal(Q.from_gpu(dB, ldb, k), got)


@pytest.mark.parametrize("P", [1, 0, 3])
def test_qr_tree_2d_wide(Q, P):
    """m < k: min(m, k) = m columns factored, the rest receive Qᵀ."""
    m, k, b = 300, 520, 64
    A = O.gaussian(m, k, 4321)
    dA = Q.to_gpu(A)
    Q.QR_tree_2D(dA, m, m, k, b, numPes=P)

    def split(mm, bw):
        L, _, _ = Q.plan_tsqr(mm, bw, P)
        return np.diff(offsets(mm, L))

    want = O.QR_tree_2D(A, b, split=None if P == 1 else split)
    assert rel(Q.from_gpu(dA, m, k), want) < 1e-11


def test_tree_entry_points_reject_bad_arguments(Q):
    """Status 1 for bad leading dimensions or ranks, 2 for b > 128; a
    rejected construct_Q1 leaves Q1 as it was."""
    from candmc_amd._lib import CQRError
    m, b, P, k = 1000, 50, 3, 60
    print("tree entry points: bad arguments", flush=True)
    f = factor(Q, O.gaussian(m, b, 31), b, P)
    dQ1 = torch.full((m * k,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(CQRError) as e:
        Q.construct_Q1(f["dY"], m - 1, f["dtau"], dQ1, m, m, b, k, numPes=P, tree_data=f["td"])
    assert e.value.status == 1
    assert bool((dQ1 == 7.0).all())
    dB = Q.to_gpu(O.gaussian(m, 5, 32))
    with pytest.raises(CQRError) as e:
        Q.apply_tsqr_QT(f["dY"], m, f["dtau"], dB, m - 1, m, b, 5, numPes=P, tree_data=f["td"])
    assert e.value.status == 1
    with pytest.raises(CQRError) as e:
        Q.tree_data_export(f["td"], m, b, P, f["L"])
    assert e.value.status == 1
    dA = Q.to_gpu(O.gaussian(400, 300, 33))
    with pytest.raises(CQRError) as e:
        Q.QR_tree_2D(dA, 400, 400, 300, 129, numPes=P)
    assert e.value.status == 2


# ---------------------------------------------------------------- no tree_data
@pytest.mark.parametrize("m,b,P,output_Y", [(8192, 64, 4, 1), (8192, 64, 4, 0), (3000, 20, 3, 1), (1024, 128, 1, 1),
                                            (1024, 128, 1, 0), (640, 64, 5, 1)])
def test_bitree_tsqr_without_tree_data(Q, m, b, P, output_Y):
    """bitree_tsqr(..., tree_data = NULL), the call of test_bitree_tsqr.cxx:319:
    per-rank dgeqrf blocks, with output_Y the senders' top triangles hold the
    nodes' V2 and rank 0's the final R; tau is rank 0's (b doubles)."""
    print(f"bitree no tree_data m={m} b={b} P={P} output_Y={output_Y}", flush=True)
    A = O.gaussian(m, b, 900 + m + b + P)
    ranges = O.split_rows(m, b, P)
    want, Rw, tauw = O.bitree_tsqr_no_tree_data([A[s:e].copy(order="F") for s, e in ranges], b, output_Y)
    dA = Q.to_gpu(A)
    dR = torch.full((b * b,), 3.0, dtype=torch.float64, device="cuda")
    dtau = torch.full((b + 1,), 5.0, dtype=torch.float64, device="cuda")
    Q.bitree_tsqr(dA, m, dR, dtau, m, b, numPes=P, output_Y=output_Y)
    got = Q.from_gpu(dA, m, b)
    R = Q.from_gpu(dR, b, b)
    sg = np.sign(np.diag(R) / np.diag(Rw))
    assert np.all(sg == 1.0)                      # same dlarfg sign rule at every node
    assert rel(R, Rw) < 1e-12 and np.all(np.tril(R, -1) == 0.0)
    assert rel(dtau.cpu().numpy()[:b], tauw) < 1e-12 and float(dtau[b]) == 5.0
    for (s, e), W in zip(ranges, want):
        assert rel(got[s:e], W) < 1e-11, (s, e)


@pytest.mark.parametrize("m,b,k,use_id", [(1000, 40, 1000, False), (777, 31, 31, False), (512, 64, 64, True)])
def test_one_rank_construct_Q1_and_apply(Q, m, b, k, use_id):
    """One rank (construct_tsqr_Q.cxx:87-102, apply_tsqr_QT.cxx:73-80):
    Q1 from [V\\R] and tau alone, as LAPACK's dorgqr / dormqr('L','N') of
    [ID; 0]; apply_tsqr_QT as dormqr('L','T')."""
    from scipy.linalg import lapack as lp
    print(f"one-rank Q m={m} b={b} k={k} ID={use_id}", flush=True)
    A = O.gaussian(m, b, 70 + m)
    dA = Q.to_gpu(A)
    dR = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dtau = torch.zeros(b, dtype=torch.float64, device="cuda")
    Q.bitree_tsqr(dA, m, dR, dtau, m, b, numPes=1, output_Y=1)
    YR, tau = Q.from_gpu(dA, m, b), dtau.cpu().numpy()
    qr, tl = O.local_tsqr(A)
    assert rel(YR, qr) < 1e-12 and rel(tau, tl) < 1e-12
    dQ1 = torch.zeros(m * k, dtype=torch.float64, device="cuda")
    if use_id:
        ID = O.gaussian(b, k, 5)
        Q.construct_Q1(dA, m, dtau, dQ1, m, m, b, k, numPes=1, ID=Q.to_gpu(ID))
        C = np.zeros((m, k), order="F")
        C[:b] = ID
        want, _, info = lp.dormqr("L", "N", qr, tl, C, max(1, k) * 64)
    else:
        Q.construct_Q1(dA, m, dtau, dQ1, m, m, b, k, numPes=1)
        Qf, _, info = lp.dorgqr(np.asfortranarray(np.hstack([qr, np.zeros((m, k - b))])), tl)
        want = Qf[:, :k]
    assert info == 0
    assert rel(Q.from_gpu(dQ1, m, k), want) < 1e-12
    B = O.gaussian(m, 9, 6)
    dB = Q.to_gpu(B)
    Q.apply_tsqr_QT(dA, m, dtau, dB, m, m, b, 9, numPes=1)
    want, _, info = lp.dormqr("L", "T", qr, tl, np.asfortranarray(B), 9 * 64)
    assert rel(Q.from_gpu(dB, m, 9), want) < 1e-12


def test_compute_invT_from_Y_two_Y(Q):
    """qr_2d.cxx:76-106: T (b1×b2, ld lda_T) := Y1ᵀ·Y2; rows past b1 untouched."""
    mb, b1, b2, ldt = 3001, 70, 130, 80
    Y1, Y2 = O.gaussian(mb, b1, 1), O.gaussian(mb, b2, 2)
    dT = torch.full((ldt * b2,), -1.0, dtype=torch.float64, device="cuda")
    Q.compute_invT_from_Y2(Q.to_gpu(Y1), mb, Q.to_gpu(Y2), mb, dT, ldt, mb, b1, b2)
    T = dT.cpu().numpy().reshape(b2, ldt).T
    assert rel(T[:b1], Y1.T @ Y2) < 1e-13
    assert np.all(T[b1:] == -1.0)
    Q.compute_invT_from_Y2(Q.to_gpu(Y1), mb, Q.to_gpu(Y2), mb, dT, ldt, 0, b1, b2)
    assert np.all(dT.cpu().numpy().reshape(b2, ldt).T[:b1] == 0.0)
