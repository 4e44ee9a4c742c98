"""Python mirror of CANDMC's alg/QR interface over the HIP C-ABI.

Function names, argument order and meaning follow the reference headers
(alg/QR/tsqr/bitree_tsqr.h, alg/QR/hh_recon/hh_recon.h, alg/QR/qr_2d/qr_2d.h).
Matrices are torch float64 tensors on the GPU holding column-major data
with the stated leading dimension (flat or any shape; only the storage is
used).  The MPI arguments of the reference (myRank, numPes, root, req_id,
cdt, pv) are accepted; on one GPU the process column collapses onto
`numPes` virtual leaves of the TSQR tree.  A non-zero status raises
CQRError (the reference asserts or aborts instead).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import CQRError, load

__all__ = [
    "Context", "context", "to_gpu", "from_gpu",
    "local_tsqr", "tree_tsqr", "bitree_tsqr", "construct_Q1", "signed_NLU", "recursive_NLU",
    "hh_recon_qr", "compute_invT_from_W", "compute_invT_from_Y", "upd_A", "update_A",
    "QR_2D", "QR_2D_2D", "tau_recon", "pack_upper", "unpack_upper", "pack_lower", "unpack_lower",
    "tree_data_size", "plan_tsqr", "apply_tsqr_QT", "QR_tree_2D", "tree_data_export", "compute_invT_from_Y2",
]


class Context:
    """One cqr_ctx: device, stream (torch's current stream) and workspace."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        s = self.lib.cqr_create(ctypes.byref(h), device)
        if s != 0:
            raise CQRError(s, "cqr_create failed")
        self.h = h
        self.device = device
        self.use_stream(torch.cuda.current_stream(device))

    def use_stream(self, stream):
        self.stream = stream
        self.check(self.lib.cqr_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)))

    def check(self, status):
        if status != 0:
            raise CQRError(status, self.lib.cqr_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.cqr_sync(self.h))

    def done(self, status, sync=True):
        """Raise on a failed launch; with sync, wait for the stream and raise
        if a TSQR tree or construct_Q1 wait gave up (cqr_sync reports it)."""
        self.check(status)
        if sync:
            self.sync()

    def last_prec(self):
        """Factorization precision of the last lstsq's X (PREC_*)."""
        return int(self.lib.cqr_lstsq_last_prec(self.h))

    def set_lstsq_escalate(self, on: bool):
        self.check(self.lib.cqr_set_lstsq_escalate(self.h, 1 if on else 0))

    def set_leaf_rows(self, rows: int):
        self.check(self.lib.cqr_set_leaf_rows(self.h, int(rows)))

    def set_lookahead(self, on: bool):
        self.check(self.lib.cqr_set_lookahead(self.h, 1 if on else 0))

    def set_wait_limit(self, polls: int):
        """Diagnostic bound on in-launch waits (0: default)."""
        self.check(self.lib.cqr_set_wait_limit(self.h, int(polls)))

    def set_check_finite(self, on: bool):
        self.check(self.lib.cqr_set_check_finite(self.h, 1 if on else 0))

    PROFILE_KINDS = ("gemm_yta", "gemm_update", "gemm_small", "splitk_reduce", "leaf_qr", "tree",
                     "hr_core", "bgemm")

    def profile(self, on: bool):
        self.check(self.lib.cqr_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self.check(self.lib.cqr_profile_reset(self.h))

    def profile_read(self):
        """{kind: (ms, flops, bytes, launches)} summed since the last reset."""
        out = {}
        for k, name in enumerate(self.PROFILE_KINDS):
            ms, fl, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
            self.check(self.lib.cqr_profile_read(self.h, k, ctypes.byref(ms), ctypes.byref(fl),
                                                 ctypes.byref(by), ctypes.byref(n)))
            out[name] = (ms.value, fl.value, by.value, n.value)
        return out

    def workspace_bytes(self) -> int:
        return int(self.lib.cqr_workspace_bytes(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.cqr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX = {}


def context(device: int = 0) -> Context:
    ctx = _CTX.get(device)
    if ctx is None:
        ctx = _CTX[device] = Context(device)
    cur = torch.cuda.current_stream(device)
    if cur.cuda_stream != ctx.stream.cuda_stream:
        ctx.use_stream(cur)
    return ctx


def _p(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise TypeError("expected a GPU tensor (there is no CPU path)")
    return ctypes.c_void_p(t.data_ptr())


def _pd(t):
    if t is not None and t.dtype != torch.float64:
        raise TypeError("expected float64")
    return _p(t)


def to_gpu(a, device=0) -> torch.Tensor:
    """Host array → flat column-major float64 GPU tensor."""
    a = np.asarray(a, dtype=np.float64)
    flat = np.ascontiguousarray(a.reshape(-1, order="F")) if a.ndim == 2 else np.ascontiguousarray(a)
    return torch.from_numpy(flat).to(f"cuda:{device}")


def from_gpu(t: torch.Tensor, m: int, n: int, ld: int | None = None) -> np.ndarray:
    """Flat column-major GPU tensor → m×n host array (Fortran order)."""
    ld = m if ld is None else ld
    h = t.detach().reshape(-1)[: ld * n].cpu().numpy()
    return np.asfortranarray(h.reshape(n, ld).T[:m, :])


def plan_tsqr(m, b, nleaves=0, leaf_rows=0):
    lib = load()
    L, R, V = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    s = lib.cqr_plan_tsqr(m, b, nleaves, leaf_rows, ctypes.byref(L), ctypes.byref(R), ctypes.byref(V))
    if s != 0:
        raise CQRError(s, f"cannot plan m={m} b={b}")
    return L.value, R.value, V.value


def tree_data_size(m, b, numPes=0):
    return int(load().cqr_tree_data_doubles(m, b, numPes))


# --- alg/QR/tsqr ----------------------------------------------------------
def local_tsqr(A, R, m, b, lda_A, tau=None, sync=True):
    """bitree_tsqr.cxx:48-71: dgeqrf of one m×b block.  R is the caller's
    scratch in the reference; here it receives the compact-WY T (b×b) when
    given.  Returns tau (b)."""
    ctx = context()
    if tau is None:
        tau = torch.empty(b, dtype=torch.float64, device=A.device)
    ctx.done(sync=sync, status=ctx.lib.cqr_local_tsqr(ctx.h, _pd(A), lda_A, m, b, _pd(tau), _pd(R)))
    return tau


def tree_tsqr(A, buf, b, tau=None, sync=True):
    """bitree_tsqr.cxx:20-34: dtpqrt of [R1; R2] in a 2b×b buffer; tau
    receives the nb×b blocked T (nb = min(b, 16)).  buf is unused."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_tree_tsqr(ctx.h, _pd(A), b, _pd(tau)))


def bitree_tsqr(A, lda_A, R, tau, m, b, myRank=0, numPes=1, root=0, req_id=0, cdt=None,
                output_Y=0, tree_data=None, sync=True):
    """bitree_tsqr.cxx:91-248.  With tree_data: the library's leaves (at
    least numPes; 0 = automatic), tau holding nleaves×b doubles and tree_data
    tree_data_size(m, b, numPes).  Without: the reference's per-rank
    semantics over numPes ranks (tau: rank 0's b doubles; with output_Y the
    senders' top triangles receive the tree's V2)."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_bitree_tsqr(ctx.h, _pd(A), lda_A, _pd(R), _pd(tau), m, b, numPes,
                                      int(output_Y), _pd(tree_data)))


def construct_Q1(Y, lda_Y, tau, Q1, lda_Q, m, b, k, myRank=0, numPes=1, root=0, cdt=None,
                 tree_data=None, ID=None, sync=True):
    """construct_tsqr_Q.cxx:34-242: Q1 (m×k) := Q·[ID; 0] (ID b×k, ld b), or
    the first k columns of the TSQR's Q when ID is None."""
    ctx = context()
    if ID is None and k == b:
        ctx.done(sync=sync, status=ctx.lib.cqr_construct_Q1(ctx.h, _pd(Y), lda_Y, _pd(tau), _pd(Q1), lda_Q, m, b, k,
                                           numPes, _pd(tree_data)))
    else:
        ctx.done(sync=sync, status=ctx.lib.cqr_construct_Q1_id(ctx.h, _pd(Y), lda_Y, _pd(tau), _pd(Q1), lda_Q, m, b,
                                              k, numPes, _pd(tree_data), _p(ID), b))


def apply_tsqr_QT(Y, lda_Y, tau, B, lda_B, m, b, k, myRank=0, numPes=1, root=0, cdt=None,
                  tree_data=None, trans=True, sync=True):
    """apply_tsqr_QT.cxx:31-165: B (m×k) := Qᵀ·B with the implicit Q of
    bitree_tsqr (trans=False applies Q)."""
    ctx = context()
    fn = ctx.lib.cqr_apply_tsqr_QT if trans else ctx.lib.cqr_apply_tsqr_Q
    ctx.check(fn(ctx.h, _pd(Y), lda_Y, _p(tau), _pd(B), lda_B, m, b, k, numPes, _pd(tree_data)))


def QR_tree_2D(A, lda_A, m, k, b, numPes=0, stop_at=-1, sync=True):
    """qr_tree_2d.cxx:30-124 over numPes virtual row blocks per panel."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_qr_tree_2d(ctx.h, _pd(A), lda_A, m, k, b, stop_at, numPes))


def tree_data_export(tree_data, m, b, numPes, rank):
    """Virtual rank `rank`'s tree_data in the layout of bitree_tsqr.cxx:213-217
    (per receiving level: dtpqrt T, min(b,16)×b, then packed upper V2)."""
    ctx = context()
    n = ctypes.c_int64()
    ctx.check(ctx.lib.cqr_tree_data_export(ctx.h, None, m, b, numPes, rank, None, ctypes.byref(n)))
    out = torch.zeros(max(n.value, 1), dtype=torch.float64, device=tree_data.device)
    if n.value:
        ctx.check(ctx.lib.cqr_tree_data_export(ctx.h, _pd(tree_data), m, b, numPes, rank, _pd(out), None))
    return out[: n.value]


# --- alg/QR/hh_recon -------------------------------------------------------
def signed_NLU(A, R, b, lda_A, lda_R, signs=None, sync=True):
    """hh_recon.cxx:43-70; signs is an int64 GPU tensor or None."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_signed_nlu(ctx.h, _pd(A), lda_A, _pd(R), lda_R, b, _p(signs)))


def recursive_NLU(A, R, b, lda_A, lda_R, signs=None, sync=True):
    """hh_recon.cxx:80-111."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_recursive_nlu(ctx.h, _pd(A), lda_A, _pd(R), lda_R, b, _p(signs)))


def hh_recon_qr(A, lda_A, m, b, W, myRank=0, numPes=1, root=0, req_id=0, cdt=None,
                T=None, invT_T=None, sync=True):
    """hh_recon.cxx:129-321: A := [Y\\R], W := −T·Y1ᵀ; numPes virtual leaves."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_hh_recon_qr_ex(ctx.h, _pd(A), lda_A, m, b, _pd(W), _pd(T), _pd(invT_T),
                                         numPes))


def compute_invT_from_W(W, b, invT, sync=True):
    """hh_recon.cxx:26-31: invT holds Y1 on entry, T⁻ᵀ on exit."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_compute_invT_from_W(ctx.h, _pd(W), b, _pd(invT)))


# --- alg/QR/qr_2d ------------------------------------------------------------
def compute_invT_from_Y(Y, lda_Y, T, lda_T, mb, b, pv=None, sync=True):
    """qr_2d.cxx:22-60."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_compute_invT_from_Y(ctx.h, _pd(Y), lda_Y, _pd(T), lda_T, mb, b))


def compute_invT_from_Y2(Y1, lda_Y1, Y2, lda_Y2, T, lda_T, mb, b1, b2, pv=None, sync=True):
    """qr_2d.cxx:76-106 (the two-Y form): T (b1×b2) := Y1ᵀ·Y2."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_compute_invT_from_Y2(ctx.h, _pd(Y1), lda_Y1, _pd(Y2), lda_Y2, _pd(T),
                                                                lda_T, mb, b1, b2))


def upd_A(Ybuf, lda_Y, A, lda_A, mb, kb, b, W, pv=None, W_is_T=False, sync=True):
    """qr_2d.cxx:224-282 (Ybuf explicit)."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_upd_A(ctx.h, _pd(Ybuf), lda_Y, _pd(A), lda_A, mb, kb, b, _pd(W),
                                1 if W_is_T else 0))


def update_A(Y, lda_Y, A, lda_A, m, k, b, W, pv=None, aggreg_Y=None, lda_aY=0, W_is_T=False, sync=True):
    """qr_2d.cxx:124-177 (Y is the panel [Y\\R])."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_update_A(ctx.h, _pd(Y), lda_Y, _pd(A), lda_A, m, k, b, _pd(W),
                                   _pd(aggreg_Y), lda_aY, 1 if W_is_T else 0))


def QR_2D(A, lda_A, m, k, b, pv=None, aggreg_Y=None, lda_aY=0, sync=True):
    """qr_2d.cxx:299-363 on one GPU."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_qr_2d(ctx.h, _pd(A), lda_A, m, k, b, _pd(aggreg_Y), lda_aY))


def QR_2D_2D(A, lda_A, m, k, b, b_sub, pv=None, aggreg_Y=None, lda_aY=0, sync=True):
    """qr_2d.cxx:772-897 on one GPU."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_qr_2d_2d(ctx.h, _pd(A), lda_A, m, k, b, b_sub, _pd(aggreg_Y), lda_aY))


# --- least-squares path ------------------------------------------------------
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2


def geqrf2d(A, lda, m, k, b, Tt=None, prec=PREC_F64, sync=True):
    """QR_2D that also returns the panel factors Tᵀ (⌈k/b⌉ blocks of b·b)
    and runs the trailing update in `prec`.  Returns Tt."""
    ctx = context()
    if Tt is None:
        Tt = torch.zeros(((k + b - 1) // b) * b * b, dtype=torch.float64, device=A.device)
    ctx.done(sync=sync, status=ctx.lib.cqr_geqrf2d(ctx.h, _pd(A), lda, m, k, b, _pd(Tt), int(prec)))
    return Tt


def apply_qt(YR, ldy, Tt, m, k, b, B, ldb, nrhs, trans=True, sync=True):
    """B := Qᵀ·B (trans) or Q·B, Q from geqrf2d (apply_tsqr_QT.cxx:31-165 analogue)."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_apply_qt(ctx.h, _pd(YR), ldy, _pd(Tt), m, k, b, _pd(B), ldb, nrhs,
                                   1 if trans else 0))


def trsm_r(YR, ldy, n, B, ldb, nrhs, trans=False, sync=True):
    """B(0:n) := R⁻¹·B (or R⁻ᵀ·B), R the upper triangle of YR."""
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_trsm_r(ctx.h, _pd(YR), ldy, n, _pd(B), ldb, nrhs, 1 if trans else 0))


def lstsq(A, lda, m, n, B, ldb, nrhs, X=None, ldx=None, b=128, prec=PREC_F64, max_iter=10, tol=1e-14,
          raise_noconverge=True):
    """min ‖A·X − B‖ with refinement on FP64 residuals; returns (X, iters,
    last_update).  A refinement that ends above tol raises CQRError with
    status 6 (CQR_ERR_NOCONVERGE) unless raise_noconverge is False; then
    the result is (X, iters, last_update, status).  context().last_prec()
    tells whether a mixed solve was refactored in FP64."""
    ctx = context()
    ldx = n if ldx is None else ldx
    if X is None:
        X = torch.zeros(ldx * nrhs, dtype=torch.float64, device=A.device)
    it, upd = ctypes.c_int(), ctypes.c_double()
    st = ctx.lib.cqr_lstsq(ctx.h, _pd(A), lda, m, n, _pd(B), ldb, nrhs, _pd(X), ldx, b, int(prec),
                           int(max_iter), float(tol), ctypes.byref(it), ctypes.byref(upd))
    if st == 6 and not raise_noconverge:
        ctx.sync()
        return X, it.value, upd.value, st
    ctx.done(st)
    return (X, it.value, upd.value) if raise_noconverge else (X, it.value, upd.value, st)


# --- batched tall-skinny QR ----------------------------------------------------
class Batch:
    r"""Many independent m_p×n_p problems (n_p ≤ 128), packed column-major one
    after another; each is factored as one hh_recon_qr panel ([Y\R] in
    place, Tᵀ in a second packed buffer)."""

    def __init__(self, m, n):
        self.ctx = context()
        self.m = np.asarray(m, dtype=np.int64)
        self.n = np.asarray(n, dtype=np.int64)
        assert self.m.shape == self.n.shape and self.m.ndim == 1
        self.np = len(self.m)
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.cqr_batch_create(self.ctx.h, ctypes.byref(h), self.np,
                                                     self.m.ctypes.data, self.n.ctypes.data))
        self.h = h
        self.aoff = np.zeros(self.np, dtype=np.int64)
        self.toff = np.zeros(self.np, dtype=np.int64)
        ta, tt = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(self.ctx.lib.cqr_batch_offsets(self.h, self.aoff.ctypes.data, self.toff.ctypes.data,
                                                      ctypes.byref(ta), ctypes.byref(tt)))
        self.total_a, self.total_t = ta.value, tt.value

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.cqr_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def groups(self):
        out = []
        for g in range(self.ctx.lib.cqr_batch_groups(self.h)):
            v = [ctypes.c_int64() for _ in range(5)]
            self.ctx.check(self.ctx.lib.cqr_batch_group_info(self.h, g, *[ctypes.byref(x) for x in v]))
            out.append(dict(zip(("b", "nprob", "leaf_rows", "leaves", "nodes"), [x.value for x in v])))
        return out

    def set_leaf_rows(self, group, rows):
        self.ctx.check(self.ctx.lib.cqr_batch_set_leaf_rows(self.ctx.h, self.h, group, rows))

    def pack(self, mats, dtype=torch.float64, device=0):
        buf = np.empty(self.total_a, dtype=np.float64)
        for p, a in enumerate(mats):
            a = np.asarray(a, dtype=np.float64)
            buf[self.aoff[p]: self.aoff[p] + a.size] = a.reshape(-1, order="F")
        return torch.from_numpy(buf).to(device=f"cuda:{device}", dtype=dtype)

    def unpack(self, A):
        h = A.detach().double().cpu().numpy()
        return [np.asfortranarray(h[self.aoff[p]: self.aoff[p] + self.m[p] * self.n[p]]
                                  .reshape(self.n[p], self.m[p]).T) for p in range(self.np)]

    def unpack_t(self, Tt):
        h = Tt.detach().double().cpu().numpy()
        return [np.asfortranarray(h[self.toff[p]: self.toff[p] + self.n[p] ** 2]
                                  .reshape(self.n[p], self.n[p]).T) for p in range(self.np)]

    def new_t(self, A):
        return torch.zeros(self.total_t, dtype=A.dtype, device=A.device)

    def qr(self, A, Tt, use_graph=False, sync=True):
        dt = {torch.float64: 0, torch.float32: 1}[A.dtype]
        assert Tt.dtype == A.dtype
        self.ctx.check(self.ctx.lib.cqr_batch_qr(self.ctx.h, self.h, _p(A), _p(Tt), dt, 1 if use_graph else 0))
        if sync:
            self.ctx.sync()

    def qr_loop(self, A, Tt, sync=True):
        self.ctx.done(self.ctx.lib.cqr_batch_qr_loop(self.ctx.h, self.h, _pd(A), _pd(Tt)), sync)

    def tune(self, A, reps=2):
        self.ctx.check(self.ctx.lib.cqr_batch_tune(self.ctx.h, self.h, _pd(A), reps))
        return self.groups()


# --- least-squares request server ---------------------------------------------
TIER_F64, TIER_MIXED = 0, 1


class _Request(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("nrhs", ctypes.c_int64),
                ("X", ctypes.c_void_p), ("ldx", ctypes.c_int64), ("tier", ctypes.c_int)]


class _Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int), ("iters", ctypes.c_int), ("slot", ctypes.c_int),
                ("prec", ctypes.c_int), ("batch", ctypes.c_int), ("last_update", ctypes.c_double), ("t_submit_ns", ctypes.c_int64), ("t_start_ns", ctypes.c_int64),
                ("t_end_ns", ctypes.c_int64), ("bytes", ctypes.c_int64)]


def _res(r):
    return {k: getattr(r, k) for k, _ in _Result._fields_}


def lstsq_workspace_bytes(m, n, nrhs, b=128):
    return int(load().cqr_lstsq_workspace_bytes(m, n, nrhs, b))


def _requests(reqs):
    arr = (_Request * len(reqs))()
    for i, (A, lda, m, n, B, ldb, nrhs, X, ldx) in enumerate(reqs):
        arr[i] = _Request(_pd(A), lda, m, n, _pd(B), ldb, nrhs, _pd(X), ldx, TIER_F64)
    return arr


def lstsq_batched(reqs, max_iter=4, tol=1e-12):
    """Least squares of many one-panel problems (n ≤ 128): reqs is a list of
    (A, lda, m, n, B, ldb, nrhs, X, ldx); one batched FP64 factorization,
    then each solve refined to tol.  Returns [(status, iters, last_update)]."""
    ctx = context()
    k = len(reqs)
    arr = _requests(reqs)
    st, it, up = (ctypes.c_int * k)(), (ctypes.c_int * k)(), (ctypes.c_double * k)()
    ctx.check(ctx.lib.cqr_lstsq_batched(ctx.h, k, arr, int(max_iter), float(tol), st, it, up))
    return [(st[i], it[i], up[i]) for i in range(k)]


def lstsq_batched_workspace_bytes(reqs):
    return int(load().cqr_lstsq_batched_workspace_bytes(len(reqs), _requests(reqs)))


class Server:
    """Least-squares request server: `nslots` slots (own stream and arena
    each) admit queued requests FIFO when their footprint fits the budget."""

    def __init__(self, nslots=4, hbm_budget=32 << 30, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        st = self.lib.cqr_server_create(ctypes.byref(h), device, nslots, hbm_budget)
        if st:
            raise CQRError(st, "server_create")
        self.h = h
        self._keep = {}

    def submit(self, A, lda, m, n, B, ldb, nrhs, X, ldx, tier=TIER_MIXED):
        r = _Request(_pd(A), lda, m, n, _pd(B), ldb, nrhs, _pd(X), ldx, int(tier))
        rid = int(self.lib.cqr_server_submit(self.h, ctypes.byref(r)))
        if rid < 0:
            raise CQRError(-rid, "server_submit")
        self._keep[rid] = (A, B, X)      # tensors stay alive until completion
        return rid

    def wait(self, rid):
        r = _Result()
        st = self.lib.cqr_server_wait(self.h, rid, ctypes.byref(r))
        if st:
            raise CQRError(st, "server_wait")
        self._keep.pop(rid, None)
        return _res(r)

    def drain(self, max_results=1024):
        arr = (_Result * max_results)()
        cnt = ctypes.c_int()
        st = self.lib.cqr_server_drain(self.h, arr, max_results, ctypes.byref(cnt))
        if st:
            raise CQRError(st, "server_drain")
        out = [_res(arr[i]) for i in range(cnt.value)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def set_batching(self, max_batch):
        st = self.lib.cqr_server_set_batching(self.h, int(max_batch))
        if st:
            raise CQRError(st, "server_set_batching")

    def pending(self):
        q, b = ctypes.c_int64(), ctypes.c_size_t()
        self.lib.cqr_server_pending(self.h, ctypes.byref(q), ctypes.byref(b))
        return q.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.cqr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- alg/shared/util ---------------------------------------------------------
def tau_recon(structure, ncol, nrow, lda, Y, tau, sync=True):
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_tau_recon(ctx.h, structure.encode()[:1], ncol, nrow, lda, _pd(Y), _pd(tau)))


def pack_upper(R_in, R_out, b, lda_in, sync=True):
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_pack_upper(ctx.h, _pd(R_in), _pd(R_out), b, lda_in))


def unpack_upper(R_in, R_out, b, lda_out, sync=True):
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_unpack_upper(ctx.h, _pd(R_in), _pd(R_out), b, lda_out))


def pack_lower(R_in, R_out, b, lda_in, has_diag, sync=True):
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_pack_lower(ctx.h, _pd(R_in), _pd(R_out), b, lda_in, int(has_diag)))


def unpack_lower(R_in, R_out, b, lda_out, has_diag, sync=True):
    ctx = context()
    ctx.done(sync=sync, status=ctx.lib.cqr_unpack_lower(ctx.h, _pd(R_in), _pd(R_out), b, lda_out, int(has_diag)))
