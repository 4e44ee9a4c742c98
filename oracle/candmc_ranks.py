"""ORACLE — the reference's QR_2D in its own process layout, on host cores.

TEST INFRASTRUCTURE ONLY: ``bench.py``'s CPU-baseline leg times it and
``tests/test_oracle.py`` checks it; nothing in ``candmc_amd`` imports it.

The reference runs ``QR_2D`` (alg/QR/qr_2d/qr_2d.cxx:299-363) on a P×1
process grid: every MPI rank owns the b-row blocks g ≡ rank (mod P) of the
matrix (2D block-cyclic with one process column, qr_2d.cxx:140-147) and
calls LAPACK single-threaded.  Here each rank is an OS process with a
single-threaded BLAS/LAPACK, its blocks kept in shared memory, and the MPI
messages become writes to shared buffers between barriers:

  panel j, root = owner of block j (the panel's top rows):
    leaf dgeqrf of the rank's active blocks      bitree_tsqr.cxx:156
    binary tree over the working ranks, root
    first (n → ⌈n/2⌉, pairs (z, z + ⌈n/2⌉)),
    dtpqrt per receiver                          bitree_tsqr.cxx:172-232
    construct_Q1: identity at the root, dtpmqrt
    down the tree, dormqr at the leaves          construct_tsqr_Q.cxx:34-242
    signed LU of Q1_top − S on the root,
    U, L, W and R broadcast                      hh_recon.cxx:267-308
    Y = Q1·U⁻¹ on every rank                     hh_recon.cxx:294-301
    T⁻ᵀ from W (every rank), Z_r = Y_rᵀ·A_r,
    Σ_r Z_r (reduce-scatter by column slices,
    then every rank reads the sum), A_r −= Y_r·(Tᵀ·Z)   qr_2d.cxx:179-282

Each step calls the same restated functions as ``candmc_oracle``
(local_tsqr, tree_tsqr, recursive_NLU, compute_invT_from_W).
"""
from __future__ import annotations

import multiprocessing as mp
import time
from multiprocessing import shared_memory

import numpy as np

TAU_BLK = 16


def _blocks_of(m, b, P):
    """Global b-row blocks of each rank (block-cyclic over P ranks)."""
    nblk = (m + b - 1) // b
    return [[g for g in range(nblk) if g % P == r] for r in range(P)]


def _levels(n):
    out = []
    while n > 1:
        out.append(n)
        n = n // 2 + n % 2
    return out


def _pairs(n):
    return [(z, z + (n + 1) // 2) for z in range(n // 2)]


class _Shared:
    """Named float64 arrays in one shared-memory segment."""

    def __init__(self, shapes, name=None):
        self.shapes = shapes
        self.offs, tot = {}, 0
        for k, shp in shapes.items():
            self.offs[k] = tot
            tot += int(np.prod(shp)) * 8
        self.shm = shared_memory.SharedMemory(create=name is None, size=max(tot, 8), name=name)
        self.arr = {k: np.ndarray(shp, dtype=np.float64, buffer=self.shm.buf, offset=self.offs[k], order="F")
                    for k, shp in shapes.items()}

    def close(self):
        self.arr = {}
        self.shm.close()


def _rank_main(r, P, m, n, b, npan, shapes, name, barrier_obj, out_q):
    from threadpoolctl import threadpool_limits
    from scipy.linalg import lapack as lp
    from scipy.linalg import solve_triangular

    from oracle import candmc_oracle as O

    class _B:   # every wait bounded: a rank that died breaks the barrier for all
        def wait(self):
            barrier_obj.wait(timeout=600)

    barrier = _B()
    with threadpool_limits(1):
        sh = _Shared(shapes, name)
        A = sh.arr[f"A{r}"]
        Rb, Cb, Zp, Zs = sh.arr["R"], sh.arr["C"], sh.arr["Zp"], sh.arr["Zs"]
        Ub, Lb, Wb, Rf = sh.arr["U"], sh.arr["L"], sh.arr["W"], sh.arr["Rf"]
        blocks = _blocks_of(m, b, P)
        mine = blocks[r]
        nblk = (m + b - 1) // b
        barrier.wait()
        t0 = time.perf_counter()
        for j in range(npan):
            c0 = j * b
            root = j % P
            # the rank's active blocks are those with global index ≥ j
            kloc = next((i for i, g in enumerate(mine) if g >= j), len(mine))
            ls = kloc * b
            act = sorted({g % P for g in range(j, nblk)}, key=lambda x: (x - root) % P)
            me = act.index(r) if r in act else -1
            if me >= 0:                                            # leaf dgeqrf
                y, tau = O.local_tsqr(A[ls:, c0:c0 + b])
                A[ls:, c0:c0 + b] = y
                Rb[r] = np.triu(y[:b, :b])
            barrier.wait()
            recs = []
            for nn in _levels(len(act)):                           # tree, root first
                for z, s in _pairs(nn):
                    if z == me:
                        Rn, V2, T = O.tree_tsqr(Rb[r], Rb[act[s]], b)
                        Rb[r] = Rn
                        recs.append((act[s], T, V2))
                barrier.wait()
            if me == 0:
                Cb[r] = np.eye(b)
            for nn in reversed(_levels(len(act))):                 # construct_Q1 down the tree
                for z, s in _pairs(nn):
                    if z == me:
                        snd, T, V2 = recs.pop()
                        a, bb, info = lp.dtpmqrt(b, V2, T, np.asfortranarray(Cb[r]), np.zeros((b, b), order="F"),
                                                 side="L", trans="N")
                        assert info == 0
                        Cb[r] = a
                        Cb[snd] = np.triu(bb)
                barrier.wait()
            Q1 = None
            if me >= 0:                                            # leaf expansion (dormqr)
                Q1 = np.zeros((A.shape[0] - ls, b), order="F")
                Q1[:b, :] = Cb[r]
                Q1 = O._dormqr("N", A[ls:, c0:c0 + b], tau, Q1)
            if r == root:                                          # signed LU of Q1_top − S
                top = np.asfortranarray(Q1[:b, :b].copy())
                Rs = np.asfortranarray(Rb[r].copy())
                signs = np.zeros(b, dtype=np.int64)
                O.recursive_NLU(top, Rs, signs)
                U = np.triu(top)
                W = U.copy()
                for i in range(b):
                    if signs[i] == -1:
                        W[: i + 1, i] *= -1.0
                Ub[:] = U
                Lb[:] = np.tril(top, -1) + np.eye(b)
                Wb[:] = W
                Rf[:] = np.triu(Rs)
            barrier.wait()
            Yu = None
            if me >= 0:                                            # Y = Q1·U⁻¹, [Y\R] into A
                Y = solve_triangular(Ub.T, Q1.T, lower=True).T
                if r == root:
                    Y[:b, :] = Lb
                    A[ls:ls + b, c0:c0 + b] = np.triu(Rf) + np.tril(Lb, -1)
                    A[ls + b:, c0:c0 + b] = Y[b:]
                else:
                    A[ls:, c0:c0 + b] = Y
                Yu = Y
            k = n - c0 - b
            if k > 0:                                              # upd_A
                invT_T = O.compute_invT_from_W(Wb, Lb)
                Zp[r][:, :k] = Yu.T @ A[ls:, c0 + b:] if Yu is not None else 0.0
                barrier.wait()
                lo, hi = r * k // P, (r + 1) * k // P
                Zs[:, lo:hi] = sum(Zp[p][:, lo:hi] for p in range(P))
                barrier.wait()
                if Yu is not None:
                    X = solve_triangular(invT_T, Zs[:, :k], lower=True)
                    A[ls:, c0 + b:] -= Yu @ X
            barrier.wait()
        out_q.put((r, time.perf_counter() - t0))
        sh.close()


def qr2d_ranks(Aglob, b, P, max_panels=None, gather=True):
    """QR_2D of the m×n matrix Aglob over P single-threaded rank processes.

    Returns (YR or None, seconds): the [Y\\R] result in global row order
    (with gather) and the longest rank's time over the panels (excluding
    process start-up and data distribution).  n must be a multiple of b.
    """
    m, n = Aglob.shape
    assert n % b == 0 and m % b == 0 and m >= n, "needs m >= n, both multiples of b"
    npan = n // b if max_panels is None else min(max_panels, n // b)
    blocks = _blocks_of(m, b, P)
    rows = [sum(min(b, m - g * b) for g in bl) for bl in blocks]
    shapes = {f"A{r}": (max(rows[r], 1), n) for r in range(P)}
    shapes.update({"R": (P, b, b), "C": (P, b, b), "Zp": (P, b, max(n - b, 1)), "Zs": (b, max(n - b, 1)),
                   "U": (b, b), "L": (b, b), "W": (b, b), "Rf": (b, b)})
    sh = _Shared(shapes)
    try:
        for r in range(P):
            o = 0
            for g in blocks[r]:
                h = min(b, m - g * b)
                sh.arr[f"A{r}"][o:o + h, :] = Aglob[g * b:g * b + h, :]
                o += h
        ctx = mp.get_context("spawn")
        barrier = ctx.Barrier(P)
        q = ctx.Queue()
        procs = [ctx.Process(target=_rank_main, args=(r, P, m, n, b, npan, shapes, sh.shm.name, barrier, q))
                 for r in range(P)]
        for p in procs:
            p.start()
        times = []
        try:
            while len(times) < P:
                try:
                    times.append(q.get(timeout=1.0))
                except Exception:       # queue.Empty: check that every rank still runs
                    dead = [p for p in procs if p.exitcode not in (None, 0)]
                    if dead:
                        raise RuntimeError(f"rank process exited with {dead[0].exitcode}")
            for p in procs:
                p.join()
                if p.exitcode != 0:
                    raise RuntimeError(f"rank process exited with {p.exitcode}")
        finally:
            for p in procs:
                if p.is_alive():
                    p.terminate()
                    p.join()
        YR = None
        if gather:
            YR = np.zeros((m, n), order="F")
            for r in range(P):
                o = 0
                for g in blocks[r]:
                    h = min(b, m - g * b)
                    YR[g * b:g * b + h, :] = sh.arr[f"A{r}"][o:o + h, :]
                    o += h
        return YR, max(t for _, t in times)
    finally:
        sh.close()
        sh.shm.unlink()
