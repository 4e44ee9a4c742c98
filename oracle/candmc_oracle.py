"""ORACLE — CPU restatement of CANDMC's communication-avoiding QR hot path.

TEST INFRASTRUCTURE ONLY.  The HIP product path's results are checked
against this module, and ``bench.py`` times it as the CPU baseline; nothing
in ``candmc_amd`` imports it, links it or calls it.

What it restates (paths relative to the CANDMC source tree):

  local_tsqr           alg/QR/tsqr/bitree_tsqr.cxx:48-71
  tree_tsqr            alg/QR/tsqr/bitree_tsqr.cxx:20-34
  bitree_tsqr          alg/QR/tsqr/bitree_tsqr.cxx:91-248
  construct_Q1         alg/QR/tsqr/construct_tsqr_Q.cxx:34-242
  construct_Q1_full    alg/QR/tsqr/construct_tsqr_Q.cxx:34-242 (any k, ID)
  apply_tsqr_QT        alg/QR/tsqr/apply_tsqr_QT.cxx:31-165
  apply_tsqr_Q         alg/QR/tsqr/construct_tsqr_Q.cxx:146-233 (general start)
  QR_tree_2D           alg/QR/qr_2d/qr_tree_2d.cxx:30-124 (one process column)
  tree_data_packed     alg/QR/tsqr/bitree_tsqr.cxx:213-217 (layout)
  signed_NLU           alg/QR/hh_recon/hh_recon.cxx:43-70
  recursive_NLU        alg/QR/hh_recon/hh_recon.cxx:80-111
  compute_invT_from_W  alg/QR/hh_recon/hh_recon.cxx:26-31
  hh_recon_qr          alg/QR/hh_recon/hh_recon.cxx:129-321 (default branch)
  compute_invT_from_Y  alg/QR/qr_2d/qr_2d.cxx:22-60
  update_A / upd_A     alg/QR/qr_2d/qr_2d.cxx:124-177, 224-282
  QR_2D                alg/QR/qr_2d/qr_2d.cxx:299-363 (1x1 process grid)
  QR_2D_2D             alg/QR/qr_2d/qr_2d.cxx:772-897 (1x1 process grid)
  tau_recon            alg/shared/util.cxx:437-465
  pack/unpack helpers  alg/shared/util.cxx:263-435

The MPI messages of the reference become plain array hand-offs between
"virtual ranks" (a Python list of per-rank blocks).  The LAPACK/BLAS
routines the reference calls are called here through scipy's LAPACK
(``dgeqrf``, ``dtpqrt``, ``dtpmqrt``, ``dormqr``, ``dorgqr``) and numpy
BLAS, with the reference's arguments.

How each function is pinned (tests/test_oracle.py):
  * bitree_tsqr / local_tsqr / tree_tsqr: R against LAPACK dgeqrf of the
    whole matrix (upper triangle up to row signs), node outputs against
    the closed form of a triangle-on-triangle QR, and against the golden
    fixtures in tests/golden.
  * construct_Q1: Q1 against dorgqr up to column signs, Q1ᵀQ1 = I,
    Q1·R = A.
  * apply_tsqr_QT / apply_tsqr_Q: against a dense m×m Q assembled from
    dorgqr of every leaf and the node transforms I − [I;V2]·T·[I;V2]ᵀ
    embedded at the leaves' top rows; Qᵀ·A = [R; 0]; Q·(Qᵀ·B) = B.
  * construct_Q1_full: equals construct_Q1 for k = b, dorgqr for one rank,
    the dense Q's first k columns, and Q·ID for ID.
  * QR_tree_2D: one rank equals LAPACK dgeqrf + dormqr; the [A, A]
    invariant of test/QR/test_qr_tree_2d.cxx:83-143 (the right half becomes
    the left half's R) for one and several ranks.
  * tree_data_packed: element positions against pack_upper and dtpqrt's T.
  * signed_NLU / recursive_NLU: L·U = Q1_top − S with the sign rule, and
    a brute-force pure-Python LU on tiny inputs.
  * hh_recon_qr: [Y\\R] equals LAPACK dgeqrf's [V\\R] (the identity that
    Householder reconstruction satisfies for m > b), W = −T·Y1ᵀ with
    T = dlarft(Y, tau) and T from LAPACK dgeqrt(nb=b).
  * compute_invT_from_W vs compute_invT_from_Y: two independent closed
    forms of T⁻ᵀ that must agree.
  * upd_A: closed form A − Y·Tᵀ·Yᵀ·A and the explicit reflector product.
  * QR_2D / QR_2D_2D: [Y\\R] equals dgeqrf(A) for m > k; the reference's
    own acceptance metric (column-relative ‖A−QR‖ ≤ 1e-9).
  * dlarft: I − V·T·Vᵀ equals the explicit product H_1···H_b.
  * drand48_elem: glibc's srand48/drand48.
No function here is left without such a check.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack as _lp
from scipy.linalg import solve_triangular as _trsv

TAU_BLK = 16   # alg/QR/tsqr/bitree_tsqr.h:6-8
BSIZE = 64     # alg/QR/hh_recon/hh_recon.cxx:13


# --------------------------------------------------------------------------
# Inputs
# --------------------------------------------------------------------------
_A48 = np.uint64(0x5DEECE66D)
_C48 = np.uint64(0xB)
_M48 = np.uint64((1 << 48) - 1)


def drand48_elem(seed):
    """First drand48() after srand48(seed), element-wise (POSIX closed form).

    srand48 puts the low 32 bits of the seed in the high 32 bits of the
    48-bit state with low bits 0x330E; one LCG step x' = a·x + c mod 2^48
    follows and the value is x'/2^48.
    """
    s = np.asarray(seed, dtype=np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    x0 = (s << np.uint64(16)) | np.uint64(0x330E)
    with np.errstate(over="ignore"):
        x1 = (_A48 * x0 + _C48) & _M48
    return x1.astype(np.float64) / float(1 << 48)


def hh_recon_test_matrix(m, b, seed_offset=99900):
    """whole_A of test/QR/test_hh_recon.cxx:78-85 (column-major m×b)."""
    row = np.arange(m, dtype=np.int64)[:, None]
    col = np.arange(b, dtype=np.int64)[None, :]
    return np.asfortranarray(drand48_elem(seed_offset + (row + col * m) * 61))


def qr_2d_test_matrix(m, k, seed_offset=99900):
    """whole_A of test/QR/test_qr_2d.cxx:96-103 (drand48() − 0.5)."""
    row = np.arange(m, dtype=np.int64)[:, None]
    col = np.arange(k, dtype=np.int64)[None, :]
    return np.asfortranarray(drand48_elem(seed_offset + (row + col * m) * 61) - 0.5)


def gaussian(m, n, seed):
    """Seeded iid N(0,1) matrix in column-major order (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.asfortranarray(rng.standard_normal((n, m)).T)


# --------------------------------------------------------------------------
# Packed-triangle layouts (alg/shared/util.cxx:263-435, util.h:686)
# --------------------------------------------------------------------------
def psz_upr(b):
    return b * (b + 1) // 2


def pack_upper(R, b):
    """util.cxx:346-358: column j contributes rows 0..j."""
    return np.concatenate([R[: j + 1, j] for j in range(b)]).copy()


def unpack_upper(p, b):
    """util.cxx:394-406 into a zero b×b buffer."""
    R = np.zeros((b, b), order="F")
    o = 0
    for j in range(b):
        R[: j + 1, j] = p[o:o + j + 1]
        o += j + 1
    return R


def pack_lower(T, b, has_diag=1):
    """util.cxx:322-335."""
    return np.concatenate([T[j + 1 - has_diag:, j] for j in range(b)]).copy()


def unpack_lower(p, b, has_diag=1):
    """util.cxx:370-383 into a zero b×b buffer."""
    T = np.zeros((b, b), order="F")
    o = 0
    for j in range(b):
        n = b - j - 1 + has_diag
        T[j + 1 - has_diag:, j] = p[o:o + n]
        o += n
    return T


def tau_recon(structure, Y, ncol, nrow):
    """util.cxx:437-465: τ_i = 2/(‖y_i‖²+1), with 2 mapped to 0."""
    tau = np.empty(ncol)
    for i in range(ncol):
        if structure == "L":
            nrm = np.linalg.norm(Y[i + 1:nrow, i])
        elif structure == "U":
            nrm = np.linalg.norm(Y[: i + 1, i])
        else:
            nrm = np.linalg.norm(Y[:nrow, i])
        t = 2.0 / (nrm * nrm + 1.0)
        tau[i] = 0.0 if t == 2.0 else t
    return tau


# --------------------------------------------------------------------------
# LAPACK-level helpers
# --------------------------------------------------------------------------
def dlarft(V, tau):
    """T of H_1···H_k = I − V·T·Vᵀ (forward, columnwise), V unit lower.

    Restates the definition LAPACK dlarft implements:
    T[i,i] = τ_i, T[0:i,i] = −τ_i·T[0:i,0:i]·V[:,0:i]ᵀ·v_i.
    """
    m, k = V.shape
    Vu = np.tril(V, -1)
    Vu[np.arange(k), np.arange(k)] = 1.0
    T = np.zeros((k, k), order="F")
    for i in range(k):
        T[i, i] = tau[i]
        if i:
            z = Vu[:, :i].T @ Vu[:, i]
            T[:i, i] = -tau[i] * (T[:i, :i] @ z)
    return T


def unit_lower(YR, b):
    """Unit-lower-trapezoidal Y from an m×b [Y\\R] buffer (qr_2d.cxx:157-162)."""
    Y = np.tril(np.asarray(YR)[:, :b], -1).copy(order="F")
    Y[np.arange(b), np.arange(b)] = 1.0
    return Y


# --------------------------------------------------------------------------
# TSQR (alg/QR/tsqr)
# --------------------------------------------------------------------------
def local_tsqr(A):
    """bitree_tsqr.cxx:48-71: LAPACK dgeqrf of one block, [V\\R] and τ."""
    qr, tau, work, info = _lp.dgeqrf(np.asfortranarray(A, dtype=np.float64))
    assert info == 0
    return np.asfortranarray(qr), tau


def tree_tsqr(R1, R2, b):
    """bitree_tsqr.cxx:20-34: dtpqrt(b, b, l=b, nb=min(b,16)) of [R1; R2].

    Returns (R, V2, T) where V2 is b×b upper triangular and T is the
    nb×b blocked triangular factor, exactly the data the reference stores.
    """
    nb = min(b, TAU_BLK)
    a, bb, t, info = _lp.dtpqrt(b, nb, np.asfortranarray(np.triu(R1)),
                                np.asfortranarray(np.triu(R2)))
    assert info == 0
    return np.triu(a), np.triu(bb), np.asfortranarray(t)


def rank_rows(m, b, P, rank, root=0):
    """bitree_tsqr.cxx:110-111: rows owned by a rank (b-row blocks)."""
    mb = ((m // b) // P) * b
    if (m // b) % P > (rank + P - root) % P:
        mb += b
    return mb


def split_rows(m, b, P):
    """Contiguous per-rank row ranges with root 0 (test_hh_recon layout).

    The reference distributes whole b-row blocks (bitree_tsqr.cxx:110-111),
    so with P > 1 it silently drops m % b trailing rows; refuse that case.
    """
    if P == 1:
        return [(0, m)]
    assert m % b == 0, "the reference's rank layout needs m % b == 0 when P > 1"
    out, o = [], 0
    for r in range(P):
        mb = rank_rows(m, b, P, r)
        out.append((o, o + mb))
        o += mb
    return out


def bitree_levels(np_work):
    """The np sequence of the tree loop, bitree_tsqr.cxx:172."""
    lv, n = [], np_work
    while n > 1:
        lv.append(n)
        n = n // 2 + n % 2
    return lv


def bitree_pairs(n):
    """(receiver, sender) pairs of one tree level, bitree_tsqr.cxx:175-208."""
    return [(r, r + (n + 1) // 2) for r in range(n // 2)]


def bitree_tsqr(blocks, b):
    """bitree_tsqr.cxx:91-248 over virtual ranks (root 0, tree_data branch).

    blocks: list of per-rank mb×b arrays.  Returns (Ys, R, taus, tree_data)
    with Ys the per-rank [V\\R] blocks, R the root's b×b R, taus per rank
    and tree_data[r] the list of (T, V2) the rank stored as a receiver, in
    level order.
    """
    P = len(blocks)
    Ys, taus, Rs = [], [], []
    for A in blocks:                          # :156-157
        y, t = local_tsqr(A)
        Ys.append(y)
        taus.append(t)
        Rs.append(np.triu(y[:b, :b]))
    tree_data = [[] for _ in range(P)]
    for n in bitree_levels(P):                # :172
        for r, s in bitree_pairs(n):          # receivers stack [R_mine; R_recv]
            Rn, V2, T = tree_tsqr(Rs[r], Rs[s], b)
            tree_data[r].append((T, V2))      # :215-217
            Rs[r] = Rn
    return Ys, Rs[0].copy(order="F"), taus, tree_data


def bitree_tsqr_no_tree_data(blocks, b, output_Y=1):
    """bitree_tsqr.cxx:91-248 without tree_data (the branch test_bitree_tsqr
    exercises, :319).  One working rank: dgeqrf of the whole block, A holds
    [V\\R] only with output_Y (:129-146).  Several: each rank's block gets
    its dgeqrf [V\\R] (:156); a receiver factors the stacked [R_mine; R_recv]
    with dgeqrf (local_tsqr, :211); with output_Y it writes the new R into its
    own top triangle (:223-224) and sends the packed upper lower half (the
    node's V2) back, which the sender unpacks into its top triangle
    (:184-189).  Returns (blocks after the call, R, rank 0's tau)."""
    P = len(blocks)
    out = [np.asfortranarray(B.copy()) for B in blocks]
    if P == 1:
        y, t = local_tsqr(blocks[0])
        if output_Y:
            out[0] = y
        return out, np.triu(y[:b, :b]), t
    taus, Rs = [], []
    for r in range(P):
        y, t = local_tsqr(blocks[r])
        out[r] = y
        taus.append(t)
        Rs.append(np.triu(y[:b, :b]))
    for n in bitree_levels(P):
        for r, s in bitree_pairs(n):
            y, _ = local_tsqr(np.vstack([Rs[r], Rs[s]]))
            Rs[r] = np.triu(y[:b, :b])
            if output_Y:
                top = out[r][:b, :b]
                top[np.triu_indices(b)] = Rs[r][np.triu_indices(b)]
                V2 = np.triu(y[b:, :b])
                top = out[s][:b, :b]
                top[np.triu_indices(b)] = V2[np.triu_indices(b)]
    return out, Rs[0].copy(order="F"), taus[0]


def construct_Q1(Ys, taus, tree_data, b, k=None):
    """construct_tsqr_Q.cxx:34-242 (tree_data branch, ID = NULL).

    Walks the tree from the root (identity at the top b×b of the root) to
    the leaves with the stored (T, V2), then applies each rank's leaf
    reflectors in TAU_BLK column blocks (dormqr, :229-232).
    """
    k = b if k is None else k
    P = len(Ys)
    if P == 1:                                # :87-107: dorgqr fast path
        q, work, info = _lp.dorgqr(np.asfortranarray(Ys[0][:, :k]), taus[0])
        assert info == 0
        return [np.asfortranarray(q)]
    Cs = [np.zeros((b, k), order="F") for _ in range(P)]
    Cs[0][:, :] = np.eye(b, k)                # :118-122
    lvls = bitree_levels(P)
    ptr = [len(td) for td in tree_data]
    for n in reversed(lvls):                  # :147-217
        for r, s in bitree_pairs(n):
            ptr[r] -= 1
            T, V2 = tree_data[r][ptr[r]]
            top = np.asfortranarray(Cs[r].copy())
            bot = np.zeros((b, k), order="F")
            # dtprfb loop of :195-208 == dtpmqrt('L','N') with the same T
            a, bb, info = _lp.dtpmqrt(b, V2, T, top, bot, side="L", trans="N")
            assert info == 0
            Cs[r] = np.asfortranarray(a)
            Cs[s] = np.asfortranarray(np.triu(bb) if k == b else bb)
    out = []
    for r in range(P):                        # :219-233
        mb = Ys[r].shape[0]
        Q = np.zeros((mb, k), order="F")
        Q[:b, :] = Cs[r]
        nb = TAU_BLK
        for j in range((b // nb) * nb, -1, -nb):
            vecs = min(b - j, nb)
            if vecs <= 0:
                continue
            c = np.asfortranarray(Q[j:, j:])
            cq, work, info = _lp.dormqr("L", "N", np.asfortranarray(Ys[r][j:, j:j + vecs]),
                                        taus[r][j:j + vecs], c, max(1, c.shape[1]) * 64)
            assert info == 0
            Q[j:, j:] = cq
        out.append(Q)
    return out


def _dormqr(trans, Y, tau, C):
    """dormqr('L', trans) of the reflectors in Y (columns with τ) on C."""
    C = np.asfortranarray(np.array(C, dtype=np.float64))
    if C.shape[1] == 0:
        return C
    c, work, info = _lp.dormqr("L", trans, np.asfortranarray(Y), tau, C, max(1, C.shape[1]) * 64)
    assert info == 0
    return np.asfortranarray(c)


def apply_tsqr_QT(Ys, taus, tree_data, Bs, b):
    """apply_tsqr_QT.cxx:31-165 over virtual ranks (root 0, tree_data branch).

    Every rank applies its leaf reflectors transposed (dlarft + dlarfb
    ('L','T'), :87-91, the same product as dormqr('L','T')); then the tree
    levels run forward and each receiver r applies dtpmqrt('L','T') with its
    stored (T, V2) to [top b rows of r; top b rows of its sender] (:131-140).
    Bs: per-rank mb×k blocks; returns the transformed blocks.
    """
    P = len(Ys)
    out = [_dormqr("T", Ys[r][:, :b], taus[r], Bs[r]) for r in range(P)]
    if P == 1:                                # :73-80
        return out
    ptr = [0] * P
    for n in bitree_levels(P):                # :104
        for r, s in bitree_pairs(n):
            T, V2 = tree_data[r][ptr[r]]
            ptr[r] += 1
            a, bb, info = _lp.dtpmqrt(b, V2, T, np.asfortranarray(out[r][:b]),
                                      np.asfortranarray(out[s][:b]), side="L", trans="T")
            assert info == 0
            out[r][:b] = a
            out[s][:b] = bb
    return out


def apply_tsqr_Q(Ys, taus, tree_data, Bs, b):
    """Q·B for the same Q: the tree levels backward with dtpmqrt('L','N')
    (the dtprfb loop of construct_tsqr_Q.cxx:195-208 on an arbitrary
    [top; bottom]), then every rank's leaf reflectors (dormqr('L','N'),
    :219-233)."""
    P = len(Ys)
    out = [np.asfortranarray(np.array(B, dtype=np.float64)) for B in Bs]
    if P > 1:
        ptr = [len(td) for td in tree_data]
        for n in reversed(bitree_levels(P)):
            for r, s in bitree_pairs(n):
                ptr[r] -= 1
                T, V2 = tree_data[r][ptr[r]]
                a, bb, info = _lp.dtpmqrt(b, V2, T, np.asfortranarray(out[r][:b]),
                                          np.asfortranarray(out[s][:b]), side="L", trans="N")
                assert info == 0
                out[r][:b] = a
                out[s][:b] = bb
    return [_dormqr("N", Ys[r][:, :b], taus[r], out[r]) for r in range(P)]


def construct_Q1_full(Ys, taus, tree_data, b, k, ID=None):
    """construct_tsqr_Q.cxx:34-242 for any k and ID: Q·[ID; 0] (ID b×k on
    the root's top rows, :123-126), or the first k columns of Q (identity
    columns at their global rows; for k ≤ b this is construct_Q1, for one
    rank the dorgqr of :87-102)."""
    sizes = [Y.shape[0] for Y in Ys]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    Es = [np.zeros((mb, k), order="F") for mb in sizes]
    if ID is not None:
        Es[0][:b, :] = ID
    else:
        for j in range(k):
            r = int(np.searchsorted(offs, j, side="right") - 1)
            Es[r][j - offs[r], j] = 1.0
    return apply_tsqr_Q(Ys, taus, tree_data, Es, b)


def QR_tree_2D(A, b, split=None, stop_at=-1):
    """qr_tree_2d.cxx:30-124 with the panels' TSQR over virtual ranks.

    split(m, b) gives the per-rank row counts of an m-row panel (None: one
    rank, i.e. dgeqrf of the panel and dormqr('L','T') of the trailing
    columns).  Per panel: bitree_tsqr with output_Y (A := leaf [V\\R]
    blocks, :58-60), the root's R over the top of the panel (:61-62),
    apply_tsqr_QT of the trailing columns (:96-97), then the next panel b
    rows and b columns further (:116-120).
    """
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    m, k = A.shape
    stop = min(m, k) if stop_at < 0 else min(m, k, stop_at)
    j = 0
    while j < stop:
        mj, bw = m - j, min(b, stop - j)
        sizes = [mj] if split is None else list(split(mj, bw))
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
        blocks = [A[j + offs[r]:j + offs[r + 1], j:j + bw] for r in range(len(sizes))]
        Ys, R, taus, td = bitree_tsqr(blocks, bw)
        for r in range(len(sizes)):
            A[j + offs[r]:j + offs[r + 1], j:j + bw] = Ys[r]
        A[j:j + bw, j:j + bw] = np.triu(R) + np.tril(A[j:j + bw, j:j + bw], -1)
        if k - j - bw > 0:
            Bs = [A[j + offs[r]:j + offs[r + 1], j + bw:] for r in range(len(sizes))]
            out = apply_tsqr_QT(Ys, taus, td, Bs, bw)
            for r in range(len(sizes)):
                A[j + offs[r]:j + offs[r + 1], j + bw:] = out[r]
        j += b
    return A


def tree_data_packed(tree_data_r, b):
    """One rank's tree_data as bitree_tsqr.cxx:213-217 stores it: per
    receiving level the dtpqrt T (min(b,16)×b, column-major) and then
    pack_upper(V2)."""
    parts = []
    for T, V2 in tree_data_r:
        parts.append(np.asarray(T).reshape(-1, order="F"))
        parts.append(pack_upper(V2, b))
    return np.concatenate(parts) if parts else np.zeros(0)


# --------------------------------------------------------------------------
# Householder reconstruction (alg/QR/hh_recon)
# --------------------------------------------------------------------------
def signed_NLU(A, R, signs):
    """hh_recon.cxx:43-70, in place on the b×b views A (L\\U) and R."""
    b = A.shape[0]
    for i in range(b):
        signs[i] = -1 if A[i, i] > 0 else 1          # :52-56
        A[i, i] -= signs[i]
        if signs[i] == -1:
            R[i, i:] *= -1.0                          # :58
        A[i + 1:, i] = A[i + 1:, i] / A[i, i]        # :61-63
        A[i + 1:, i + 1:] -= np.outer(A[i + 1:, i], A[i, i + 1:])   # :64-66 dger


def recursive_NLU(A, R, signs):
    """hh_recon.cxx:80-111 (recursion at b/2 until b ≤ BSIZE)."""
    b = A.shape[0]
    if b <= BSIZE:
        signed_NLU(A, R, signs)
        return
    b1 = b // 2
    recursive_NLU(A[:b1, :b1], R[:b1, :b1], signs[:b1])
    for i in range(b1):
        if signs[i] == -1:
            R[i, b1:] *= -1.0                          # :99-102
    A[:b1, b1:] = _trsv(A[:b1, :b1], A[:b1, b1:], lower=True, unit_diagonal=True)   # :105
    A[b1:, :b1] = _trsv(A[:b1, :b1].T, A[b1:, :b1].T, lower=True).T                 # :106 (R,U,N,N)
    A[b1:, b1:] -= A[b1:, :b1] @ A[:b1, b1:]                                         # :108
    recursive_NLU(A[b1:, b1:], R[b1:, b1:], signs[b1:])


def hh_recon_qr(A, b, P=1):
    """hh_recon.cxx:129-321 with P virtual ranks (root 0).

    A is the global m×b panel (column-major).  Returns (YR, W, extra):
    YR is the m×b result the ranks leave in A ([Y\\R], Y unit lower with
    implicit diagonal, R = S·R_tsqr on top), W = U·S (b×b upper), and extra
    holds the intermediate R, Q1, signs and the TSQR data for white-box
    checks.
    """
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    m = A.shape[0]
    ranges = split_rows(m, b, P)
    blocks = [A[s:e, :b].copy(order="F") for s, e in ranges]
    Ys, R, taus, tree_data = bitree_tsqr(blocks, b)                 # :235
    Q1s = construct_Q1(Ys, taus, tree_data, b)                       # :267
    Q1 = Q1s[0]
    signs = np.zeros(b, dtype=np.int64)
    R_s = R.copy(order="F")
    recursive_NLU(Q1[:b, :b], R_s, signs)                            # :282
    W = np.triu(Q1[:b, :b]).copy(order="F")                          # :284
    U = W.copy()
    out = []
    for r, (s, e) in enumerate(ranges):
        Yr = Ys[r].copy(order="F")
        if r == 0:
            Yr[:b, :b] = np.triu(R_s)                                # :287 copy_upper(R, A)
            Q = Q1s[0]
            if Q.shape[0] > b:                                       # :299
                Q[b:, :] = _trsv(U.T, Q[b:, :].T, lower=True).T
            low = np.tril(Q, -1)                                     # :301 copy_lower
            Yr = np.triu(Yr) + low
        else:
            Yr = _trsv(U.T, Q1s[r].T, lower=True).T                  # :294
        out.append(Yr)
    YR = np.asfortranarray(np.vstack(out))
    for i in range(b):                                               # :304-308
        if signs[i] == -1:
            W[: i + 1, i] *= -1.0
    extra = dict(R_tsqr=R, signs=signs, U=U, taus=taus, tree_data=tree_data,
                 Q1=[q for q in Q1s], leaf_Y=Ys)
    return YR, W, extra


def compute_invT_from_W(W, Y1):
    """hh_recon.cxx:26-31: solve Wᵀ·X = −Y1 for X = T⁻ᵀ (lower b×b)."""
    return _trsv(np.triu(W).T, -np.asarray(Y1), lower=True)


def compute_invT_from_Y(Y):
    """qr_2d.cxx:22-60: T⁻ᵀ = tril(YᵀY) with the diagonal halved."""
    S = np.tril(Y.T @ Y)
    S[np.diag_indices_from(S)] *= 0.5
    return S


def T_from_W(W, Y1):
    """T such that Q = I − Y·T·Yᵀ, from W = −T·Y1ᵀ (upper b×b)."""
    return np.triu(np.linalg.inv(compute_invT_from_W(W, Y1)).T)


def upd_A(Y, A, invT_T):
    """qr_2d.cxx:224-282 on one rank: A := A − Y·(T⁻ᵀ)⁻¹·(Yᵀ·A).

    Y is the unit-lower-trapezoidal panel (mb×b), invT_T the lower b×b T⁻ᵀ.
    """
    YT = Y.T @ A                                                     # :259
    YT = _trsv(invT_T, YT, lower=True)                               # :271
    return A - Y @ YT                                                # :275


def QR_2D(A, b, P_panel=1, max_panels=None, copy=True):
    """qr_2d.cxx:299-363 on a 1×1 process grid.

    Right-looking: per panel hh_recon_qr on min(k,b) columns (P_panel
    virtual ranks), then update_A with T⁻ᵀ from W (:179-208).  Returns the
    m×k [Y\\R] buffer, no τ (the reference reconstructs τ with tau_recon).
    max_panels stops after that many panels (a bounded timing sample).
    """
    A = np.asfortranarray(np.array(A, dtype=np.float64)) if copy else A
    m, k = A.shape
    j = 0
    npan = 0
    while True:
        mj, kj = m - j, k - j
        bj = min(kj, b)
        YR, W, _ = hh_recon_qr(A[j:, j:j + bj], bj, P_panel)
        A[j:, j:j + bj] = YR
        npan += 1
        if kj - b > 0 and mj - b > 0 and (max_panels is None or npan < max_panels):   # :318
            Y = unit_lower(YR, bj)
            invT_T = compute_invT_from_W(W, Y[:bj, :bj])
            Ct = A[j:, j + bj:].T                                   # row-major view of C
            YT = _trsv(invT_T, (Ct @ Y).T, lower=True)              # upd_A :259-271
            Ct -= YT.T @ Y.T                                        # :275, in place
            j += b
        else:
            break
    return A


def QR_2D_2D(A, b, b_sub):
    """qr_2d.cxx:772-897 on a 1×1 grid: QR_2D on b-wide blocks with b_sub
    panels, then one aggregated upd_A with T⁻ᵀ = compute_invT_from_Y."""
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    m, k = A.shape
    j = 0
    while True:
        mj, kj = m - j, k - j
        if kj - b > 0 and mj - b > 0:                               # :793
            A[j:, j:j + b] = QR_2D(A[j:, j:j + b], b_sub)          # :846
            Y = unit_lower(A[j:, j:j + b], b)
            A[j:, j + b:] = upd_A(Y, A[j:, j + b:], compute_invT_from_Y(Y))   # :873
            j += b
        else:
            A[j:, j:] = QR_2D(A[j:, j:], b_sub)                   # :890
            break
    return A


# --------------------------------------------------------------------------
# Least squares (no reference counterpart; pinned by scipy.linalg.lstsq)
# --------------------------------------------------------------------------
def apply_QT(YR, B, b=None):
    """Qᵀ·B for Q = H_1···H_n stored as [Y\\R] (τ from tau_recon)."""
    m, n = YR.shape
    tau = tau_recon("L", YR, n, m)
    Y = unit_lower(YR, n)
    B = np.array(B, dtype=np.float64, order="F")
    for i in range(n):
        v = Y[:, i]
        B -= tau[i] * np.outer(v, v @ B)
    return B


def conditioned_matrix(m, n, kappa, seed):
    """A = U·diag(s)·Vᵀ with U (m×n) and V (n×n) orthonormal from FP64 QR of
    seeded Gaussians and s geometric from 1 to 1/κ (the mixed-precision
    least-squares workload)."""
    U, _ = np.linalg.qr(gaussian(m, n, seed))
    V, _ = np.linalg.qr(gaussian(n, n, seed + 1))
    s = np.geomspace(1.0, 1.0 / kappa, n)
    return np.asfortranarray((U * s) @ V.T)


def lstsq_qr(A, B, b=128):
    """x = R⁻¹·(Qᵀb)[0:n] through QR_2D (FP64)."""
    YR = QR_2D(A, b)
    n = A.shape[1]
    QtB = apply_QT(YR, B)
    return _trsv(np.triu(YR[:n, :n]), QtB[:n], lower=False)


# --------------------------------------------------------------------------
# Metrics
# --------------------------------------------------------------------------
def column_relative_residual(A, YR):
    """test/QR/test_hh_recon.cxx:234-259: Σ_i ‖(A − Q·R)(:,i)‖/‖A(:,i)‖."""
    m, n = A.shape
    tau = tau_recon("L", YR, n, m)
    if m >= n:
        # Q·R with Q = H_0···H_{n−1} formed by dorgqr (same product, blocked)
        q, work, info = _lp.dorgqr(np.asfortranarray(YR.copy()), tau)
        assert info == 0
        QR = q @ np.triu(YR[:n, :n])
    else:
        R = np.zeros((m, n), order="F")
        R[:n, :] = np.triu(YR[:n, :n])
        QR = R.copy(order="F")
        Y = unit_lower(YR, n)
        for i in range(n - 1, -1, -1):
            v = Y[:, i]
            QR -= tau[i] * np.outer(v, v @ QR)
    D = A - QR
    return float(np.sum(np.sqrt(np.sum(D * D, 0) / np.sum(A * A, 0))))


def explicit_q(YR, n=None):
    """Q[:, 0:n] from [Y\\R] with τ from tau_recon (dorgqr)."""
    m, k = YR.shape
    tau = tau_recon("L", YR, k, m)
    q, work, info = _lp.dorgqr(np.asfortranarray(YR.copy()), tau)
    assert info == 0
    return q
