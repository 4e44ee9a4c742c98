"""Pins the CPU oracle (CPU only).

Every oracle function is checked against something it does not itself
compute: glibc, LAPACK fixtures frozen in tests/golden, closed forms that
Householder QR satisfies, or brute force on tiny inputs.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg import lapack as lp

from oracle import candmc_oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(np.float64).eps


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_drand48_matches_glibc():
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    seeds = [0, 1, 99900, 99900 + 61 * 12345, 2**31 - 1, 2**32 + 7, 99900 + 61 * (8191 + 63 * 8192)]
    want = []
    for s in seeds:
        libc.srand48(ctypes.c_long(s))
        want.append(libc.drand48())
    assert np.array_equal(O.drand48_elem(np.array(seeds)), np.array(want))


def test_hh_recon_test_matrix_is_fixture_input():
    g = gold("hh_recon_136x17.npz")
    assert np.array_equal(O.hh_recon_test_matrix(136, 17), g["A"])
    g = gold("qr2d_96x48.npz")
    assert np.array_equal(O.qr_2d_test_matrix(96, 48), g["A"])


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0)
    for b in (1, 5, 16):
        R = np.triu(rng.standard_normal((b, b)))
        p = O.pack_upper(R, b)
        assert p.size == O.psz_upr(b)
        assert np.array_equal(O.unpack_upper(p, b), R)
        L = np.tril(R.T)
        assert np.array_equal(O.unpack_lower(O.pack_lower(L, b, 1), b, 1), L)


@pytest.mark.parametrize("b", [1, 2, 5, 16, 33])
def test_packed_triangles_match_lapack_dtrttp(b):
    """util.cxx:322-406 store triangles column by column, LAPACK's packed
    storage: pack_upper / unpack_upper are dtrttp / dtpttr with 'U',
    pack_lower / unpack_lower(has_diag=1) with 'L'; has_diag=0 is the 'L'
    pack with each column's diagonal entry removed."""
    rng = np.random.default_rng(100 + b)
    A = np.asfortranarray(rng.standard_normal((b, b)))
    ap_u, info = lp.dtrttp(A, uplo="U")
    assert info == 0
    assert np.array_equal(O.pack_upper(A, b), ap_u)
    full_u, info = lp.dtpttr(b, ap_u, uplo="U")
    assert info == 0
    assert np.array_equal(O.unpack_upper(ap_u, b), np.triu(full_u))
    ap_l, info = lp.dtrttp(A, uplo="L")
    assert info == 0
    assert np.array_equal(O.pack_lower(A, b, 1), ap_l)
    full_l, info = lp.dtpttr(b, ap_l, uplo="L")
    assert np.array_equal(O.unpack_lower(ap_l, b, 1), np.tril(full_l))
    # strictly lower: drop entry j·b − j(j−1)/2 (the diagonal of column j)
    diag_pos = [j * b - j * (j - 1) // 2 for j in range(b)]
    strict = np.delete(ap_l, diag_pos)
    assert np.array_equal(O.pack_lower(A, b, 0), strict)
    assert np.array_equal(O.unpack_lower(strict, b, 0), np.tril(A, -1))
    # a row-major pack (the transpose's order) differs once b > 2
    if b > 2:
        assert not np.array_equal(O.pack_upper(A, b), lp.dtrttp(np.asfortranarray(A.T), uplo="L")[0])


def test_tree_data_packed_offsets_follow_index_formula():
    """bitree_tsqr.cxx:213-217: per receiving level, the dtpqrt T (nb×b,
    ld nb, nb = min(b, 16)) at offset i + j·nb, then pack_upper(V2) with
    V2(i, j), i ≤ j, at nb·b + j(j+1)/2 + i; levels follow one another."""
    for b in (5, 20):
        nb = min(b, O.TAU_BLK)
        per = nb * b + b * (b + 1) // 2
        levels = []
        for lv in range(3):
            T = np.array([[1e6 * lv + 1e3 * j + i for j in range(b)] for i in range(nb)])
            V2 = np.array([[-(1e6 * lv + 1e3 * j + i) for j in range(b)] for i in range(b)])
            levels.append((T, V2))
        p = O.tree_data_packed(levels, b)
        assert p.size == 3 * per
        for lv, (T, V2) in enumerate(levels):
            for j in range(b):
                for i in range(nb):
                    assert p[lv * per + i + j * nb] == T[i, j]
                for i in range(j + 1):
                    assert p[lv * per + nb * b + j * (j + 1) // 2 + i] == V2[i, j]


def test_dlarft_equals_reflector_product_and_dgeqrt():
    A = O.gaussian(40, 9, 3)
    qr, tau = O.local_tsqr(A)
    V = O.unit_lower(qr, 9)
    T = O.dlarft(V, tau)
    H = np.eye(40)
    for i in range(9):            # brute force: H_1 H_2 ... H_k
        v = V[:, i]
        H = H @ (np.eye(40) - tau[i] * np.outer(v, v))
    assert np.abs(np.eye(40) - V @ T @ V.T - H).max() < 1e-14
    _, t, _ = lp.dgeqrt(9, np.asfortranarray(A))
    assert np.abs(np.triu(t) - T).max() < 1e-14


def test_tau_recon_recovers_lapack_tau():
    A = O.gaussian(64, 12, 5)
    qr, tau = O.local_tsqr(A)
    assert np.abs(O.tau_recon("L", qr, 12, 64) - tau).max() < 1e-15


def test_tree_tsqr_closed_form_and_fixture():
    g = gold("tpqrt_b20.npz")
    b = 20
    R, V2, T = O.tree_tsqr(g["R1"], g["R2"], b)
    assert np.abs(R - g["R"]).max() < 1e-14 and np.abs(V2 - g["V2"]).max() < 1e-14
    # Closed form: Q = I − [I;V2]·T_full·[I;V2]ᵀ is orthogonal and Qᵀ[R1;R2] = [R;0]
    Vf = np.vstack([np.eye(b), V2])
    tau = np.array([T[i % 16, i] for i in range(b)])
    Tf = O.dlarft(Vf, tau)
    Q = np.eye(2 * b) - Vf @ Tf @ Vf.T
    assert np.abs(Q.T @ Q - np.eye(2 * b)).max() < 1e-14
    QtA = Q.T @ np.vstack([g["R1"], g["R2"]])
    assert np.abs(QtA[:b] - R).max() < 1e-13 and np.abs(QtA[b:]).max() < 1e-13
    # Same R as dgeqrf of the stacked matrix (same dlarfg sign rule)
    qr, _ = O.local_tsqr(np.vstack([g["R1"], g["R2"]]))
    assert np.abs(np.triu(qr[:b]) - R).max() < 1e-13


@pytest.mark.parametrize("m,b,P", [(136, 17, 4), (8192, 64, 4), (1000, 50, 3), (512, 32, 1), (640, 64, 5)])
def test_bitree_R_and_construct_Q1(m, b, P):
    A = O.hh_recon_test_matrix(m, b)
    ranges = O.split_rows(m, b, P)
    blocks = [A[s:e].copy(order="F") for s, e in ranges]
    Ys, R, taus, td = O.bitree_tsqr(blocks, b)
    qr, _ = O.local_tsqr(A)
    Rl = np.triu(qr[:b])
    sg = np.sign(np.diag(R) / np.diag(Rl))
    assert np.abs(R - sg[:, None] * Rl).max() / np.abs(Rl).max() < 1e-13
    Q1 = np.vstack(O.construct_Q1(Ys, taus, td, b))
    assert np.abs(Q1.T @ Q1 - np.eye(b)).max() < 1e-13
    assert np.abs(Q1 @ R - A).max() / np.abs(A).max() < 1e-13
    Ql = O.explicit_q(qr)[:, :b]
    assert np.abs(Q1 - Ql * sg[None, :]).max() < 1e-12


@pytest.mark.parametrize("m,b,P", [(256, 16, 4), (300, 12, 3), (96, 16, 1)])
def test_bitree_no_tree_data_rebuilds_A(m, b, P):
    """The no-tree_data branch leaves enough in A to rebuild it: the leaf
    reflectors below each rank's diagonal, the node reflectors [e_i; V2(:,i)]
    in the senders' top triangles, R in rank 0's.  Q·[R; 0] = A by brute
    force (explicit reflector products), R is LAPACK's up to row signs, and
    each V2 is dtpqrt's for the same node (same dlarfg sign rule)."""
    rng = np.random.default_rng(m + P)
    A = rng.standard_normal((m, b))
    ranges = O.split_rows(m, b, P)
    blocks = [A[s:e].copy(order="F") for s, e in ranges]
    out, R, tau0 = O.bitree_tsqr_no_tree_data(blocks, b, 1)
    X = np.zeros((m, b))
    X[:b] = R
    def H(rows, v):                       # I − τ·u·uᵀ on the given rows
        tau = 2.0 / (v @ v)
        E = np.eye(m)
        E[np.ix_(rows, rows)] -= tau * np.outer(v, v)
        return E
    Qm = np.eye(m)
    for r, (s0, e0) in enumerate(ranges):        # leaves: H_0 ⋯ H_{b−1}
        for i in range(b):
            v = np.concatenate([[1.0], out[r][i + 1:, i]])
            Qm = Qm @ H(list(range(s0 + i, e0)), v)
    node_q = np.eye(m)
    for n in O.bitree_levels(P):                 # levels forward: Q_tree = Q_1 Q_2 ⋯
        for z, snd in O.bitree_pairs(n):
            for i in range(b):
                v = np.concatenate([[1.0], out[snd][:i + 1, i]])
                rows = [ranges[z][0] + i] + [ranges[snd][0] + q for q in range(i + 1)]
                node_q = node_q @ H(rows, v)
    assert np.abs(Qm @ node_q @ X - A).max() < 1e-13
    qr, t = O.local_tsqr(A)
    Rl = np.triu(qr[:b])
    sg = np.sign(np.diag(R) / np.diag(Rl))
    assert np.abs(R - sg[:, None] * Rl).max() < 1e-13
    assert np.array_equal(tau0, O.local_tsqr(blocks[0])[1])
    if P > 1:
        _, _, _, td = O.bitree_tsqr(blocks, b)
        z, snd = O.bitree_pairs(P)[0]
        assert np.abs(np.triu(out[snd][:b, :b]) - td[z][0][1]).max() < 1e-13


def _brute_signed_lu(M):
    b = M.shape[0]
    A = [[float(M[i, j]) for j in range(b)] for i in range(b)]
    s = [0] * b
    for i in range(b):
        s[i] = -1 if A[i][i] > 0 else 1
        A[i][i] -= s[i]
        for r in range(i + 1, b):
            A[r][i] /= A[i][i]
        for r in range(i + 1, b):
            for c in range(i + 1, b):
                A[r][c] -= A[r][i] * A[i][c]
    return np.array(A), np.array(s)


def test_signed_nlu_brute_force():
    Q, _ = np.linalg.qr(O.gaussian(30, 6, 11))
    A = np.asfortranarray(Q[:6, :6].copy())
    R = np.triu(O.gaussian(6, 6, 12))
    R0 = R.copy()
    signs = np.zeros(6, dtype=np.int64)
    O.signed_NLU(A, R, signs)
    Ab, sb = _brute_signed_lu(Q[:6, :6])
    assert np.array_equal(signs, sb)
    assert np.abs(A - Ab).max() < 1e-14
    L = np.tril(A, -1) + np.eye(6)
    assert np.abs(L @ np.triu(A) - (Q[:6, :6] - np.diag(signs))).max() < 1e-14
    assert np.array_equal(R, np.where(signs == -1, -1.0, 1.0)[:, None] * R0)


def test_recursive_nlu_matches_unblocked():
    Q, _ = np.linalg.qr(O.gaussian(400, 128, 13))
    Ar = np.asfortranarray(Q[:128, :128].copy())
    Au = Ar.copy(order="F")
    Rr = np.triu(O.gaussian(128, 128, 14))
    Ru = Rr.copy(order="F")
    sr = np.zeros(128, dtype=np.int64)
    su = np.zeros(128, dtype=np.int64)
    O.recursive_NLU(Ar, Rr, sr)
    O.signed_NLU(Au, Ru, su)
    assert np.array_equal(sr, su)
    assert np.abs(Ar - Au).max() < 1e-12
    assert np.array_equal(Rr, Ru)
    L = np.tril(Ar, -1) + np.eye(128)
    assert np.abs(L @ np.triu(Ar) - (Q[:128, :128] - np.diag(sr))).max() < 1e-13


def test_hh_recon_matches_golden_dgeqrf():
    g = gold("hh_recon_136x17.npz")
    YR, W, _ = O.hh_recon_qr(g["A"], 17, P=4)
    assert np.abs(YR - g["YR"]).max() / np.abs(g["YR"]).max() < 1e-13
    Y = O.unit_lower(YR, 17)
    T = np.triu(g["T"])
    assert np.abs(W + T @ Y[:17, :17].T).max() < 1e-13          # W = −T·Y1ᵀ
    assert O.column_relative_residual(g["A"], YR) < 1e-9           # test_hh_recon.cxx:263


def test_hh_recon_c1_summary():
    g = gold("c1_summary.npz")
    A = O.hh_recon_test_matrix(8192, 64)
    YR, W, _ = O.hh_recon_qr(A, 64, P=4)
    R = np.triu(YR[:64, :64])
    assert np.abs(np.diag(R) - g["Rdiag"]).max() / np.abs(g["Rdiag"]).max() < 1e-13
    tau = O.tau_recon("L", YR, 64, 8192)
    assert np.abs(tau - g["tau"]).max() < 1e-13
    assert np.abs(np.linalg.norm(np.tril(YR, -1), axis=0) - g["Ycolnorm"]).max() < 1e-12


def test_invT_two_closed_forms_agree():
    A = O.gaussian(720, 40, 17)
    YR, W, _ = O.hh_recon_qr(A, 40, P=3)
    Y = O.unit_lower(YR, 40)
    a = O.compute_invT_from_W(W, Y[:40, :40])
    b = O.compute_invT_from_Y(Y)
    assert np.abs(a - b).max() / np.abs(b).max() < 1e-13


def test_upd_A_equals_reflector_product():
    A = O.gaussian(90, 30, 19)
    YR, W, _ = O.hh_recon_qr(A[:, :10], 10, P=2)
    Y = O.unit_lower(YR, 10)
    tau = O.tau_recon("L", YR, 10, 90)
    C = A[:, 10:].copy()
    out = O.upd_A(Y, C, O.compute_invT_from_W(W, Y[:10, :10]))
    ref = C.copy()
    for i in range(10):          # Qᵀ C = H_b ··· H_1 C, H_1 applied first
        v = Y[:, i]
        ref = ref - tau[i] * np.outer(v, v @ ref)
    assert np.abs(out - ref).max() < 1e-13


def leading_rows(YR, tol=1e-8):
    """Rows of R above the first pivot below tol·max|diag| (rank-revealing cut)."""
    d = np.abs(np.diag(YR))
    small = np.nonzero(d < tol * d.max())[0]
    return int(small[0]) if small.size else d.size


@pytest.mark.parametrize("name,b", [("qr2d_8x4.npz", 2), ("qr2d_96x48.npz", 8)])
def test_qr_2d_reference_inputs(name, b):
    # test_qr_2d.cxx's generator gives numerically singular matrices: the
    # reference itself only checks the column-relative residual (:292-297).
    g = gold(name)
    YR = O.QR_2D(g["A"], b)
    assert O.column_relative_residual(g["A"], YR) < 1e-9
    r = leading_rows(g["YR"])
    assert r >= 2
    assert np.abs(np.triu(YR[:r]) - np.triu(g["YR"][:r])).max() < 1e-13


def test_qr_2d_matches_golden_gaussian():
    g = gold("qr2d_gauss_160x96.npz")
    for b in (8, 16, 32):
        YR = O.QR_2D(g["A"], b)
        assert np.abs(YR - g["YR"]).max() / np.abs(g["YR"]).max() < 1e-13
        assert O.column_relative_residual(g["A"], YR) < 1e-9


def test_qr_2d_square_last_sign_quirk():
    A = O.gaussian(48, 48, 37)
    YR = O.QR_2D(A, 8)
    qr, _ = O.local_tsqr(A)
    d = np.abs(YR - qr)
    d[47, 47] = abs(YR[47, 47] + qr[47, 47])                        # test_qr_2d.cxx:135-137
    assert d.max() / np.abs(qr).max() < 1e-12
    assert abs(YR[47, 47] - qr[47, 47]) > 1e-3 * abs(qr[47, 47])


def test_qr_2d_2d_equals_qr_2d():
    A = O.gaussian(256, 96, 23)
    a = O.QR_2D(A, 16)
    b = O.QR_2D_2D(A, 32, 16)
    assert np.abs(a - b).max() / np.abs(a).max() < 1e-13


def test_lstsq_matches_scipy():
    A = O.gaussian(300, 40, 29)
    B = O.gaussian(300, 3, 31)
    x = O.lstsq_qr(A, B, b=16)
    xs = sl.lstsq(A, B)[0]
    assert np.abs(x - xs).max() < 1e-12


def test_apply_qt_against_dorgqr_and_norm():
    """Qᵀ·B from the reflectors: its top k rows are Q1ᵀ·B with Q1 from
    LAPACK dorgqr, and Qᵀ is orthogonal (norm preserved column by column)."""
    A = O.gaussian(200, 48, 37)
    B = O.gaussian(200, 3, 38)
    YR = O.QR_2D(A, 16)
    QtB = O.apply_QT(YR, B)
    Q1 = O.explicit_q(YR)
    assert np.abs(QtB[:48] - Q1.T @ B).max() < 1e-12
    assert np.allclose(np.linalg.norm(QtB, axis=0), np.linalg.norm(B, axis=0), rtol=1e-13)


def test_conditioned_matrix_has_the_stated_spectrum():
    A = O.conditioned_matrix(300, 40, 1e6, 5)
    s = np.linalg.svd(A, compute_uv=False)
    assert abs(s[0] - 1.0) < 1e-12 and abs(s[-1] * 1e6 - 1.0) < 1e-6


# ---- implicit TSQR Q (apply_tsqr_QT, construct_Q1 in full, QR_tree_2D) ----
def _dense_tsqr_q(Ys, taus, td, b):
    """Dense m×m Q = Q_leaves·Q_tree built without dormqr/dtpmqrt: each
    leaf's full orthogonal factor from dorgqr, each node as the explicit
    product of its reflectors u_i = [e_i at the receiver's top; V2(:, i) at
    the sender's top] with τ_i = diag of the blocked T, nodes in level
    order."""
    sizes = [Y.shape[0] for Y in Ys]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    m = offs[-1]
    Q = np.zeros((m, m))
    for r, Y in enumerate(Ys):
        mb = sizes[r]
        a = np.zeros((mb, mb), order="F")
        a[:, :b] = Y[:, :b]
        q, _, info = lp.dorgqr(a, taus[r])
        assert info == 0
        Q[offs[r]:offs[r + 1], offs[r]:offs[r + 1]] = q
    ptr = [0] * len(Ys)
    nb = min(b, O.TAU_BLK)
    for n in O.bitree_levels(len(Ys)):
        for r, s in O.bitree_pairs(n):
            T, V2 = td[r][ptr[r]]
            ptr[r] += 1
            G = np.eye(m)
            for i in range(b):
                u = np.zeros(m)
                u[offs[r] + i] = 1.0
                u[offs[s]:offs[s] + b] = V2[:, i]
                tau = T[i % nb, i]
                G = G - tau * np.outer(G @ u, u)
            Q = Q @ G
    return Q, offs


def _leaf_split(m, b, L):
    return [(l * m) // L for l in range(L + 1)]


@pytest.mark.parametrize("m,b,L,k", [(160, 8, 5, 7), (300, 12, 4, 20), (96, 16, 1, 5), (260, 10, 7, 33)])
def test_apply_tsqr_against_dense_q(m, b, L, k):
    rng = np.random.default_rng(m + b + L)
    A = rng.standard_normal((m, b))
    B = rng.standard_normal((m, k))
    o = _leaf_split(m, b, L)
    Ys, R, taus, td = O.bitree_tsqr([A[o[r]:o[r + 1]].copy(order="F") for r in range(L)], b)
    Q, offs = _dense_tsqr_q(Ys, taus, td, b)
    assert np.abs(Q.T @ Q - np.eye(m)).max() < 1e-13
    split = lambda X: [X[o[r]:o[r + 1]] for r in range(L)]
    qtb = np.vstack(O.apply_tsqr_QT(Ys, taus, td, split(B), b))
    assert np.abs(qtb - Q.T @ B).max() < 1e-12
    qb = np.vstack(O.apply_tsqr_Q(Ys, taus, td, split(B), b))
    assert np.abs(qb - Q @ B).max() < 1e-12
    # Qᵀ·A = [R; 0] with R at the root's top rows
    qta = np.vstack(O.apply_tsqr_QT(Ys, taus, td, split(A), b))
    want = np.zeros((m, b))
    want[:b] = R
    assert np.abs(qta - want).max() / np.abs(R).max() < 1e-13


@pytest.mark.parametrize("m,b,L", [(200, 8, 5), (120, 12, 1), (256, 16, 3)])
def test_construct_q1_full(m, b, L):
    rng = np.random.default_rng(7 * m + L)
    A = rng.standard_normal((m, b))
    o = _leaf_split(m, b, L)
    Ys, R, taus, td = O.bitree_tsqr([A[o[r]:o[r + 1]].copy(order="F") for r in range(L)], b)
    Q, _ = _dense_tsqr_q(Ys, taus, td, b)
    for k in (b, b + 5, 2 * b + 3, 3):
        q1 = np.vstack(O.construct_Q1_full(Ys, taus, td, b, k))
        assert np.abs(q1 - Q[:, :k]).max() < 1e-12
    assert np.abs(np.vstack(O.construct_Q1_full(Ys, taus, td, b, b))
                  - np.vstack(O.construct_Q1(Ys, taus, td, b))).max() < 1e-13
    ID = rng.standard_normal((b, b))
    qid = np.vstack(O.construct_Q1_full(Ys, taus, td, b, b, ID=ID))
    assert np.abs(qid - Q[:, :b] @ ID).max() < 1e-12
    if L == 1:   # one rank: dorgqr of the leaf (construct_tsqr_Q.cxx:87-102)
        k = b + 4
        a = np.zeros((m, k), order="F")
        a[:, :b] = Ys[0][:, :b]
        q, _, info = lp.dorgqr(a, taus[0])
        assert np.abs(np.vstack(O.construct_Q1_full(Ys, taus, td, b, k)) - q).max() < 1e-13


def _aa_defect(A2, k):
    """test_qr_tree_2d.cxx:83-143 metric for QR of [A, A] on one rank."""
    m = A2.shape[0]
    L, Rr = A2[:, :k], A2[:, k:2 * k]
    f = 0.0
    for c in range(k):
        for r in range(m):
            if r > c:
                f += Rr[r, c] ** 2
            else:
                d = L[r, c] - Rr[r, c]
                f += min(d * d, (d / L[r, c]) ** 2)
    return np.sqrt(f)


def test_qr_tree_2d_one_rank_is_dgeqrf_and_dormqr():
    rng = np.random.default_rng(11)
    m, k, b = 90, 40, 8
    A = rng.standard_normal((m, k))
    out = O.QR_tree_2D(A, b)
    qr, tau, _, info = lp.dgeqrf(np.asfortranarray(A.copy()))
    assert np.abs(out - qr).max() < 1e-12
    # stop_at: factor 24 columns, Qᵀ on the rest
    out = O.QR_tree_2D(A, b, stop_at=24)
    qr, tau, _, _ = lp.dgeqrf(np.asfortranarray(A[:, :24].copy()))
    rest, _, _ = lp.dormqr("L", "T", qr, tau, np.asfortranarray(A[:, 24:].copy()), 4096)
    assert np.abs(out[:, :24] - qr).max() < 1e-12
    assert np.abs(out[:, 24:] - rest).max() < 1e-12
    # the HR-based QR_2D computes the same factorization for m > k
    assert np.abs(O.QR_tree_2D(A, b) - O.QR_2D(A, b)).max() < 1e-11


@pytest.mark.parametrize("ranks", [1, 3])
def test_qr_tree_2d_aa_invariant(ranks):
    m, k, b = 64, 24, 8
    A = O.qr_2d_test_matrix(m, k)
    split = None if ranks == 1 else (lambda mm, bw: np.diff(_leaf_split(mm, bw, min(ranks, mm // bw))))
    out = O.QR_tree_2D(np.hstack([A, A]), b, split=split, stop_at=min(m, k))
    assert _aa_defect(out, k) < 1e-10
    # R (up to row signs) is LAPACK's
    qr, _, _, _ = lp.dgeqrf(np.asfortranarray(A.copy()))
    Rl, Rt = np.triu(qr[:k]), np.triu(out[:k, :k])
    sg = np.sign(np.diag(Rt) / np.diag(Rl))
    assert np.abs(Rt - sg[:, None] * Rl).max() < 1e-11


def test_tree_data_packed_layout():
    m, b, L = 96, 20, 4
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m, b))
    o = _leaf_split(m, b, L)
    Ys, R, taus, td = O.bitree_tsqr([A[o[r]:o[r + 1]].copy(order="F") for r in range(L)], b)
    nb = min(b, O.TAU_BLK)
    for r in range(L):
        p = O.tree_data_packed(td[r], b)
        assert p.size == len(td[r]) * (nb * b + b * (b + 1) // 2)
        for lv, (T, V2) in enumerate(td[r]):
            base = lv * (nb * b + b * (b + 1) // 2)
            Tb = p[base:base + nb * b].reshape(b, nb).T
            assert np.array_equal(Tb, T)
            assert np.array_equal(O.unpack_upper(p[base + nb * b:base + nb * b + b * (b + 1) // 2], b),
                                  np.triu(V2))
            # the blocked T's diagonal is the nodes' τ: the reflectors reproduce [R_r; R_s] → [R; 0]
            assert np.all(np.abs(np.diag(T[np.arange(b) % nb, np.arange(b)])) <= 2.0)


@pytest.mark.parametrize("m,n,b,P", [(1024, 256, 64, 4), (768, 192, 64, 3), (512, 256, 64, 8)])
def test_qr2d_on_rank_processes_matches_lapack(m, n, b, P):
    """The CPU baseline's restatement of QR_2D on a P×1 grid of rank
    processes (b-row blocks cyclic, tree over the working ranks, HR on the
    panel's owner, Σ Yᵀ·A reduced across ranks) gives dgeqrf's [Y\\R]; with
    P = 8 and 8 blocks the last panels run on fewer working ranks."""
    from oracle import candmc_ranks as CR
    A = O.gaussian(m, n, 21 + P)
    YR, sec = CR.qr2d_ranks(A, b, P)
    q, tau, work, info = sl.lapack.dgeqrf(np.asfortranarray(A))
    assert info == 0 and sec > 0
    assert np.abs(YR - q).max() / np.abs(q).max() < 1e-12
