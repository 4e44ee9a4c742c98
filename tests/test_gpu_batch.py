"""GPU parity of the batched tall-skinny QR: every problem of a mixed batch
(different m and n) against LAPACK dgeqrf / the oracle's HR, eager and
graph-replayed runs identical, the per-problem loop identical, FP32 buffers
within FP32 rounding, and the autotuner's choice still correct."""
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


SHAPES = [(300, 16), (1500, 16), (16, 16), (1024, 32), (2500, 32), (700, 64), (64, 64), (3000, 64),
          (256, 128), (2049, 128), (128, 128), (5000, 128), (400, 16), (900, 32)]


def problems(seed=0):
    m = [s[0] for s in SHAPES]
    n = [s[1] for s in SHAPES]
    mats = [O.gaussian(mm, nn, seed + 7 * i) for i, (mm, nn) in enumerate(SHAPES)]
    return m, n, mats


def check_against_lapack(mats, outs, tts):
    for A, YR, Tt in zip(mats, outs, tts):
        m, n = A.shape
        ref, _ = O.local_tsqr(A)                                 # dgeqrf [V\\R]
        if m == n:                                               # test_qr_2d.cxx:135-137
            YR = YR.copy(order="F")
            YR[m - 1, n - 1] *= -1.0
            assert rel(YR[:, : n - 1], ref[:, : n - 1]) < 1e-11
        else:
            assert rel(YR, ref) < 1e-11
        assert O.column_relative_residual(A, YR) < 1e-9
        T = O.dlarft(O.unit_lower(YR, n), O.tau_recon("L", YR, n, m))
        c = n - 1 if m == n else n
        assert rel(np.tril(Tt)[:c, :c], T.T[:c, :c]) < 1e-11


def test_batch_groups_and_results(Q):
    m, n, mats = problems()
    bt = Q.Batch(m, n)
    assert sorted(g["b"] for g in bt.groups()) == [16, 32, 64, 128]
    A = bt.pack(mats)
    Tt = bt.new_t(A)
    bt.qr(A, Tt)
    check_against_lapack(mats, bt.unpack(A), bt.unpack_t(Tt))


def test_batch_graph_replay_matches_eager(Q):
    m, n, mats = problems(3)
    bt = Q.Batch(m, n)
    A0 = bt.pack(mats)
    A, Tt = A0.clone(), bt.new_t(A0)
    bt.qr(A, Tt)
    eager = (A.cpu().numpy().copy(), Tt.cpu().numpy().copy())
    A2, T2 = A0.clone(), bt.new_t(A0)
    for _ in range(3):                 # first call captures, later ones replay
        A2.copy_(A0)
        torch.cuda.synchronize()
        bt.qr(A2, T2, use_graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(A2.cpu().numpy(), eager[0])
        assert np.array_equal(T2.cpu().numpy(), eager[1])


def test_batch_loop_baseline_matches(Q):
    m, n, mats = problems(5)
    bt = Q.Batch(m, n)
    A = bt.pack(mats)
    Tt = bt.new_t(A)
    bt.qr_loop(A, Tt)
    check_against_lapack(mats, bt.unpack(A), bt.unpack_t(Tt))


def test_batch_fp32_buffers(Q):
    m, n, mats = problems(9)
    bt = Q.Batch(m, n)
    A = bt.pack(mats, dtype=torch.float32)
    Tt = bt.new_t(A)
    bt.qr(A, Tt)
    for a32, YR in zip(mats, bt.unpack(A)):
        a = np.asarray(a32, dtype=np.float32).astype(np.float64)
        if a.shape[0] == a.shape[1]:
            YR = YR.copy(order="F")
            YR[-1, -1] *= -1.0
        assert O.column_relative_residual(a, YR) < 1e-4         # FP32 storage of [Y\R]


def test_batch_autotune_keeps_results(Q):
    m, n, mats = problems(11)
    bt = Q.Batch(m, n)
    A = bt.pack(mats)
    groups = bt.tune(A, reps=1)
    assert all(g["leaf_rows"] >= g["b"] for g in groups)
    Tt = bt.new_t(A)
    bt.qr(A, Tt)
    check_against_lapack(mats, bt.unpack(A), bt.unpack_t(Tt))


def test_batch_at_the_bench_distribution(Q):
    """96 problems drawn like the bench's batch (m ~ U[256, 16384], n in
    {16, 32, 64, 128}): trees of many depths share the level-by-level
    launches, and the row tiles of the Y product reach m = 16384; every
    problem against LAPACK, eager and graph-replayed."""
    rng = np.random.default_rng(13)
    m = [int(x) for x in rng.integers(256, 16385, 96)]
    n = [int(x) for x in rng.choice([16, 32, 64, 128], 96)]
    m[0], n[0] = 16384, 128                       # the largest shape of the distribution
    mats = [O.gaussian(mm, nn, 100 + i) for i, (mm, nn) in enumerate(zip(m, n))]
    bt = Q.Batch(m, n)
    A0 = bt.pack(mats)
    A, Tt = A0.clone(), bt.new_t(A0)
    bt.qr(A, Tt)
    check_against_lapack(mats, bt.unpack(A), bt.unpack_t(Tt))
    A2, T2 = A0.clone(), bt.new_t(A0)
    for _ in range(2):
        A2.copy_(A0)
        torch.cuda.synchronize()
        bt.qr(A2, T2, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(A2, A) and torch.equal(T2, Tt)


def test_batch_rejects_wide_problem(Q):
    from candmc_amd._lib import CQRError
    with pytest.raises(CQRError):
        Q.Batch([100, 300], [20, 129])
