#!/usr/bin/env python3
"""Benchmark of the MI355X CA-QR hot path.

Workload (one "step"): FP64 QR_2D of a 32768×4096 iid N(0,1) matrix with
panel width b = 128 — TSQR + Householder reconstruction panels and the
compact-WY trailing update on FP64 MFMA — on one GPU.  The input is
generated once, kept resident in HBM, and restored with a device copy
before each step (the copy is outside the per-step event timing).

Flops follow the reference's bench (bench/QR/bench_qr_2d.cxx:87):
F = 2mn² − ⅔n³, no credit for the reconstruction overhead.

Prints ONE JSON line.  Run:  python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QR GFLOP/s (% roofline) at m=32768,n=4096; LSQ solves/s + p99 at 1e-10"
PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (MFMA and vector)
PEAK_HBM_GBS = 8000.0
EPS = np.finfo(np.float64).eps


def qr_flops(m, n):
    return 2.0 * m * n * n - (2.0 / 3.0) * n ** 3


def gaussian(m, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.asfortranarray(rng.standard_normal((n, m)).T)


def apply_q(YR, v, trans):
    """Q·v or Qᵀ·v for Q = H_1···H_n stored as [Y\\R] (τ_i = 2/(1+‖y_i‖²))."""
    m, n = YR.shape
    v = v.copy()
    order = range(n) if trans else range(n - 1, -1, -1)
    for i in order:
        y = YR[i + 1:, i]
        tau = 2.0 / (1.0 + y @ y)
        w = v[i] + y @ v[i + 1:]
        v[i] -= tau * w
        v[i + 1:] -= (tau * w) * y
    return v


def accuracy(A, YR, nprobe=2, seed=7):
    """Randomized ‖A − QR‖/(‖A‖·n·ε) and ‖I − QᵀQ‖ estimates (FP64, host)."""
    m, n = A.shape
    rng = np.random.default_rng(seed)
    R = np.triu(YR[:n, :n])
    normA = np.linalg.norm(A)
    res, orth = 0.0, 0.0
    for _ in range(nprobe):
        x = rng.standard_normal(n)
        r = np.zeros(m)
        r[:n] = R @ x
        d = A @ x - apply_q(YR, r, trans=False)
        res = max(res, np.linalg.norm(d) / (normA * np.linalg.norm(x)))
        v = rng.standard_normal(m)
        w = apply_q(YR, apply_q(YR, v, trans=False), trans=True)
        orth = max(orth, np.linalg.norm(w - v) / np.linalg.norm(v))
    return res / (n * EPS), orth


def host_info():
    """The host the CPU lines ran on: usable cores, CPU model, LAPACK build."""
    from threadpoolctl import threadpool_info
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = [d for d in threadpool_info() if d.get("user_api") == "blas"]
    lapack = ", ".join(f"{d.get('internal_api')} {d.get('version')} ({os.path.basename(d.get('filepath', ''))}, "
                       f"{d.get('num_threads')} threads)" for d in blas) or "unknown"
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "cpu_model": model,
            "lapack": lapack}


def cpu_dgeqrf(A):
    """Multithreaded LAPACK dgeqrf of the whole matrix (the reference's
    sequential QR benchmark, bench/QR/bench_qr_seq.cxx:66, calls dgeqrf)."""
    from scipy.linalg import lapack
    from threadpoolctl import threadpool_info
    m, n = A.shape
    threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    W = np.asfortranarray(A.copy())
    t = time.perf_counter()
    _, _, work, info = lapack.dgeqrf(W, lwork=-1)
    qr_, tau, _, info = lapack.dgeqrf(W, lwork=int(work[0]), overwrite_a=True)
    dt = time.perf_counter() - t
    assert info == 0
    return {"value": round(qr_flops(m, n) / dt / 1e9, 2), "unit": "GFLOP/s", "cores": int(threads),
            "kind": "lapack", "sample": f"dgeqrf of the same {m}x{n} matrix", "seconds": round(dt, 3)}


def _panels_flops(m, n, b, panels):
    p = min(panels * b, n)
    return qr_flops(m, n) - (qr_flops(m - p, n - p) if p < n else 0.0)


def cpu_threaded(A, b, max_seconds):
    """The oracle's QR_2D (numpy + LAPACK) in one process on threaded BLAS."""
    from threadpoolctl import threadpool_info

    from oracle import candmc_oracle as O
    m, n = A.shape
    threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    # the full factorization when it fits the time budget, else the first
    # panels (flops = F(m,n) − F(m−pb, n−pb))
    W = np.asfortranarray(A.copy())
    t = time.perf_counter()
    O.QR_2D(W, b, max_panels=2, copy=False)
    t2 = time.perf_counter() - t
    est_full = t2 * qr_flops(m, n) / _panels_flops(m, n, b, 2) * 0.8
    panels = n // b if est_full <= max_seconds else max(2, int(max_seconds / t2 * 2))
    panels = min(panels, n // b)
    W = np.asfortranarray(A.copy())
    t = time.perf_counter()
    O.QR_2D(W, b, max_panels=panels, copy=False)
    dt = time.perf_counter() - t
    sample = (f"QR_2D of the same {m}x{n} matrix, b={b}, one process, threaded BLAS, "
              + ("full factorization" if panels >= n // b else f"first {panels} panels"))
    return {"value": round(_panels_flops(m, n, b, panels) / dt / 1e9, 2), "unit": "GFLOP/s",
            "cores": int(threads), "kind": "port", "sample": sample, "seconds": round(dt, 3)}


def cpu_baseline(A, b, max_seconds):
    """The reference's CPU path as it runs: QR_2D on a P×1 process grid, one
    single-threaded LAPACK rank per host core (P = the cores this process may
    use), restated in oracle/candmc_ranks.py; plus the same restatement in
    one process on threaded BLAS and a multithreaded dgeqrf line."""
    from oracle import candmc_ranks as CR
    m, n = A.shape
    P = len(os.sched_getaffinity(0))
    _, t2 = CR.qr2d_ranks(A, b, P, max_panels=2, gather=False)
    est_full = t2 * qr_flops(m, n) / _panels_flops(m, n, b, 2) * 0.8
    panels = n // b if est_full <= max_seconds else max(2, int(max_seconds / t2 * 2))
    panels = min(panels, n // b)
    _, dt = CR.qr2d_ranks(A, b, P, max_panels=panels, gather=False)
    sample = (f"QR_2D of the same {m}x{n} matrix, b={b}, on {P} single-threaded rank processes "
              f"(P x 1 grid, b-row blocks cyclic), "
              + ("full factorization" if panels >= n // b else f"first {panels} panels"))
    return {"value": round(_panels_flops(m, n, b, panels) / dt / 1e9, 2), "unit": "GFLOP/s", "cores": int(P),
            "kind": "port", "sample": sample, "seconds": round(dt, 3), "host": host_info(),
            "threaded_blas": cpu_threaded(A, b, max_seconds), "dgeqrf": cpu_dgeqrf(A)}


# --------------------------------------------------------------------------
# Secondary legs: mixed-precision least squares, batched QR, serving
# --------------------------------------------------------------------------
def _timed(fn, torch):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t) * 1e3


def _orthonormal(qr, torch, m, n, seed):
    """Q (m×n) with orthonormal columns: our QR of a Gaussian, Q·[I; 0]."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    G = torch.randn(n, m, generator=g, dtype=torch.float64, device="cuda").reshape(-1)
    Tt = qr.geqrf2d(G, m, m, n, 128)
    E = torch.zeros(n, m, dtype=torch.float64, device="cuda")
    E[:, :n] = torch.eye(n, dtype=torch.float64, device="cuda")
    E = E.reshape(-1)
    qr.apply_qt(G, m, Tt, m, n, 128, E, m, n, trans=False)
    return E.reshape(n, m).T            # column-major view as m×n


def lsq_leg(qr, torch, m=65536, n=2048, kappas=(1e2, 1e6, 1e8), seed=11):
    """Mixed-precision least squares: A = U·diag(s)·Vᵀ (s geometric 1..1/κ),
    b = A·x_true + 1e-6·N(0,1); FP64 solve vs FP32-update and BF16-update
    solves refined on FP64 residuals; error against the FP64 GPU solution."""
    U = _orthonormal(qr, torch, m, n, seed)
    V = _orthonormal(qr, torch, n, n, seed + 1)
    g = torch.Generator(device="cuda").manual_seed(seed + 2)
    xt = torch.randn(n, 1, generator=g, dtype=torch.float64, device="cuda")
    noise = torch.randn(m, 1, generator=g, dtype=torch.float64, device="cuda")
    out = {"m": m, "n": n, "nrhs": 1, "seed": seed, "tol_mixed": 1e-10, "cases": []}
    flops = qr_flops(m, n)
    for kappa in kappas:
        s = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device="cuda")
        A = ((U * s) @ V.T).T.contiguous().reshape(-1)        # column-major m×n
        B = ((U * s) @ (V.T @ xt) + 1e-6 * noise).reshape(-1)
        (X64, it64, u64, st64), t64 = _timed(lambda: qr.lstsq(A, m, m, n, B, m, 1, prec=qr.PREC_F64, max_iter=6,
                                                               tol=1e-12, raise_noconverge=False), torch)
        x64 = X64.double()
        case = {"kappa": kappa, "fp64_ms": round(t64, 2), "fp64_iters": it64, "fp64_status": st64,
                "fp64_last_update": float(f"{u64:.3g}")}
        ctx = qr.context()
        for name, prec in (("fp32", qr.PREC_F32), ("bf16", qr.PREC_BF16)):
            (Xm, itm, upd, st), tm = _timed(lambda: qr.lstsq(A, m, m, n, B, m, 1, prec=prec, max_iter=40,
                                                              tol=1e-10, raise_noconverge=False), torch)
            err = float((Xm - x64).abs().max() / x64.abs().max())
            used = {0: "fp64", 1: "fp32", 2: "bf16"}.get(ctx.last_prec(), "none")
            case[name] = {"ms": round(tm, 2), "iters": itm, "status": st, "factorized_in": used,
                          "last_update": float(f"{upd:.3g}"), "rel_err_vs_fp64": float(f"{err:.3g}"),
                          "met_1e-10": bool(err <= 1e-10), "speedup_vs_fp64": round(t64 / tm, 3),
                          "gflops_factor_equiv": round(flops / (tm * 1e-3) / 1e9, 1)}
        out["cases"].append(case)
        del A, B
    return out


def tree_leg(qr, torch, m=32768, b=128, k=3968, seed=19, reps=3):
    """Implicit TSQR Q (apply_tsqr_QT) of one 32768×128 panel's tree applied
    to the 32768×3968 trailing block, and the tree-updated QR_tree_2D of a
    16384×2048 matrix (one rank: dgeqrf + dormqr per panel; automatic
    leaves: TSQR + apply_tsqr_QT per panel)."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    L, _, _ = qr.plan_tsqr(m, b, 0)
    P = torch.randn(m * b, generator=g, dtype=torch.float64, device="cuda")
    R = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    tau = torch.zeros(L * b, dtype=torch.float64, device="cuda")
    td = torch.zeros(qr.tree_data_size(m, b, 0), dtype=torch.float64, device="cuda")
    qr.bitree_tsqr(P, m, R, tau, m, b, numPes=0, output_Y=1, tree_data=td)
    B = torch.randn(m * k, generator=g, dtype=torch.float64, device="cuda")
    run = lambda: qr.apply_tsqr_QT(P, m, tau, B, m, m, b, k, numPes=0, tree_data=td)
    run()
    ms = min(_timed(run, torch)[1] for _ in range(reps))
    flops = 4.0 * m * b * k + 2.0 * L * b * b * k + 6.0 * (L - 1) * b * b * k
    out = {"apply_tsqr_QT": {"m": m, "b": b, "k": k, "leaves": L, "ms": round(ms, 3),
                             "tflops_alg": round(flops / ms / 1e9, 2)}}
    mm, kk = 16384, 2048
    A0 = torch.randn(mm * kk, generator=g, dtype=torch.float64, device="cuda")
    A = torch.empty_like(A0)
    for name, P_ in (("one_rank", 1), ("leaves", 0)):
        def f():
            qr.QR_tree_2D(A, mm, mm, kk, b, numPes=P_)
        A.copy_(A0)
        f()
        best = 1e30
        for _ in range(2):
            A.copy_(A0)
            best = min(best, _timed(f, torch)[1])
        out[f"qr_tree_2d_{name}"] = {"m": mm, "k": kk, "b": b, "ms": round(best, 2),
                                     "gflops": round(qr_flops(mm, kk) / best / 1e6, 1)}
    return out


def batch_leg(qr, torch, nprob=2048, seed=13, loop=True):
    """2048 tall-skinny QRs, m ~ U[256, 16384], n ∈ {16, 32, 64, 128}:
    batched launches (autotuned leaf heights, graph-replayed) against the
    per-problem launch loop, FP64 and FP32 buffers."""
    rng = np.random.default_rng(seed)
    m = rng.integers(256, 16385, nprob)
    n = rng.choice([16, 32, 64, 128], nprob)
    bt = qr.Batch(m, n)
    g = torch.Generator(device="cuda").manual_seed(seed)
    A0 = torch.randn(bt.total_a, generator=g, dtype=torch.float64, device="cuda")
    A = A0.clone()
    Tt = bt.new_t(A)
    F = float(np.sum(2.0 * m * n.astype(np.float64) ** 2 - (2.0 / 3.0) * n.astype(np.float64) ** 3))
    byt = float(np.sum(2.0 * m * n * 8.0))
    t_roof = float(np.sum(np.maximum(2.0 * m * n * 8.0 / (PEAK_HBM_GBS * 1e9),
                                     (2.0 * m * n.astype(np.float64) ** 2) / (PEAK_FP64_TFLOPS * 1e12))))
    tuned = bt.tune(A0, reps=1)
    res = {"nprob": nprob, "seed": seed, "flops": F, "bytes": byt, "roofline_ms_fp64": round(t_roof * 1e3, 3),
           "groups": tuned}

    def timed_runs(fn, reps=3):
        ts = []
        for _ in range(reps):
            A.copy_(A0)
            _, t = _timed(fn, torch)
            ts.append(t)
        return min(ts)

    bt.qr(A, Tt)                                   # warm
    t_eager = timed_runs(lambda: bt.qr(A, Tt))
    A.copy_(A0)
    bt.qr(A, Tt, use_graph=True)                   # capture
    t_graph = timed_runs(lambda: bt.qr(A, Tt, use_graph=True))
    res["fp64"] = {"batched_ms": round(t_eager, 2), "graph_ms": round(t_graph, 2),
                   "gflops": round(F / (t_graph * 1e-3) / 1e9, 1),
                   "roofline_frac": round(t_roof * 1e3 / t_graph, 4)}
    if loop:
        t_loop = timed_runs(lambda: bt.qr_loop(A, Tt), reps=1)
        res["fp64"]["per_problem_loop_ms"] = round(t_loop, 2)
        res["fp64"]["speedup_vs_loop"] = round(t_loop / t_graph, 2)
    A32 = A0.float()
    T32 = bt.new_t(A32)
    A32w = A32.clone()
    bt.qr(A32w, T32)
    ts = []
    for _ in range(3):
        A32w.copy_(A32)
        _, t = _timed(lambda: bt.qr(A32w, T32), torch)
        ts.append(t)
    res["fp32_buffers"] = {"batched_ms": round(min(ts), 2), "note": "FP32 storage, computed in FP64"}
    bt.close()
    return res


def serve_leg(qr, torch, nreq=1000, rates=(500.0, 2000.0, 8000.0), nslots=4, seed=17, n_fixed=None):
    """Poisson streams of least-squares requests: m lognormal (median 4096,
    σ = 1, capped at 65536), n ~ U[8, 256], 1-16 right-hand sides, column
    scaling to κ ≤ 1e6.  For both tiers, with the server's batching of
    one-panel requests (n ≤ 128) off and on: p50/p99 latency and solves/s
    at each offered rate, then a burst of all requests at once (the
    capacity).  Each request's X is compared with the FP64 tier's X from
    the unbatched server (refined to 1e-12)."""
    rng = np.random.default_rng(seed)
    ms = np.minimum(65536, np.maximum(1, rng.lognormal(np.log(4096), 1.0, nreq))).astype(np.int64)
    ns = rng.integers(8, 257, nreq) if n_fixed is None else np.full(nreq, n_fixed)
    ms = np.maximum(ms, ns)
    ks = rng.integers(1, 17, nreq)
    kap = 10.0 ** rng.uniform(0, 6, nreq)
    g = torch.Generator(device="cuda").manual_seed(seed)
    reqs = []
    for m, n, k, kp in zip(ms, ns, ks, kap):
        scale = torch.logspace(0, -np.log10(kp), int(n), dtype=torch.float64, device="cuda")
        A = (torch.randn(int(n), int(m), generator=g, dtype=torch.float64, device="cuda") * scale[:, None]).reshape(-1)
        B = torch.randn(int(k) * int(m), generator=g, dtype=torch.float64, device="cuda")
        X = torch.zeros(int(n) * int(k), dtype=torch.float64, device="cuda")
        reqs.append((A, B, X, int(m), int(n), int(k)))
    torch.cuda.synchronize()
    out = {"nreq": nreq, "nslots": nslots, "seed": seed, "one_panel_requests": int(np.sum(ns <= 128)),
           "m": "lognormal(median 4096, sigma 1) capped at 65536",
           "n": "U[8,256]" if n_fixed is None else str(n_fixed), "nrhs": "U{1..16}",
           "kappa": "column scaling 1..1e-6 (log-uniform)", "tiers": {}}

    def stream(srv, tier, rate):
        """Submits all requests (Poisson arrivals at `rate`, or all at once
        for rate None) and collects them."""
        t0 = time.perf_counter()
        due = t0 + (np.cumsum(rng.exponential(1.0 / rate, nreq)) if rate else np.zeros(nreq))
        ids = []
        for (A, B, X, m, n, k), d in zip(reqs, due):
            while time.perf_counter() < d:
                pass
            ids.append(srv.submit(A, m, m, n, B, m, k, X, n, tier))
        res = [srv.wait(i) for i in ids]
        lat = np.array([(r["t_end_ns"] - r["t_submit_ns"]) * 1e-6 for r in res])
        span = (max(r["t_end_ns"] for r in res) - min(r["t_submit_ns"] for r in res)) * 1e-9
        return res, {"offered_qps": rate if rate else "burst", "solves_per_s": round(nreq / span, 1),
                     "p50_ms": round(float(np.percentile(lat, 50)), 2),
                     "p99_ms": round(float(np.percentile(lat, 99)), 2),
                     "mean_iters": round(float(np.mean([r["iters"] for r in res])), 2),
                     "mean_batch": round(float(np.mean([r["batch"] for r in res])), 2),
                     "noconverge": int(sum(r["status"] == 6 for r in res)),
                     "failed": int(sum(r["status"] not in (0, 6) for r in res))}

    ref = None
    for tier, tname in ((qr.TIER_F64, "fp64_1e-12"), (qr.TIER_MIXED, "mixed_1e-10")):
        out["tiers"][tname] = {}
        for batching in (1, 32):
            srv = qr.Server(nslots=nslots, hbm_budget=64 << 30)
            srv.set_batching(batching)
            # warm every slot's arena with the largest request
            big = int(np.argmax(ms * ns))
            A, B, X, m, n, k = reqs[big]
            for i in [srv.submit(A, m, m, n, B, m, k, X, n, tier) for _ in range(nslots)]:
                srv.wait(i)
            rows = [stream(srv, tier, rate)[1] for rate in rates]
            res, burst = stream(srv, tier, None)
            srv.close()
            xs = [X.clone() for _, _, X, _, _, _ in reqs]
            if ref is None:
                ref = xs      # the FP64 tier without batching: the reference X
            errs = np.array([float((x - r).abs().max() / r.abs().max().clamp_min(1e-300)) for x, r in zip(xs, ref)])
            burst["err_vs_fp64_max"] = float(f"{errs.max():.3g}")
            burst["err_vs_fp64_p99"] = float(f"{np.percentile(errs, 99):.3g}")
            rows.append(burst)
            out["tiers"][tname]["batching_off" if batching == 1 else "batching_32"] = rows
    out["capacity_solves_per_s"] = {f"{t}/{b}": rows[-1]["solves_per_s"]
                                    for t, d in out["tiers"].items() for b, rows in d.items()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--m", type=int, default=32768)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--b", type=int, default=128)
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-lsq", action="store_true", help="skip the mixed-precision least-squares leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched tall-skinny QR leg")
    ap.add_argument("--no-tree", action="store_true", help="skip the implicit-Q (apply_tsqr_QT) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-loop", action="store_true", help="skip the per-problem loop baseline of the batch")
    ap.add_argument("--no-qr", action="store_true",
                    help="skip the headline QR_2D (to profile the other legs on their own)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the library "
              "factors each matrix on a single MI355X and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    # Hardware queues per process, read when HIP starts (nothing here has
    # touched the GPU yet).  HIP's default of four is fewer than the
    # streams of this run (torch's, the library context's two, one per
    # server slot): streams then share queues and the slots' work
    # serialises (served capacity at n = 64: 3085 solves/s on four queues,
    # 5211 on eight).  The headline QR is the same either way.  (The
    # environment may already name four, HIP's default: raise it.)
    try:
        queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        queues = 4
    if queues < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    m, n, b = args.m, args.n, args.b
    A_host = gaussian(m, n, args.seed)
    # the CPU baseline starts its rank processes before this process touches
    # the GPU (a child forked from a process with the GPU initialised does
    # not start other programs cleanly everywhere)
    cpu = None if args.no_cpu else cpu_baseline(A_host, b, args.cpu_seconds)

    import torch

    from candmc_amd import qr

    if args.no_qr:   # profiling the other legs on their own
        out = {"metric": METRIC, "value": None, "unit": "GFLOP/s", "n_gpus": 1,
               "note": "headline QR_2D skipped (--no-qr)"}
    else:
        dev = torch.device("cuda:0")
        A0 = qr.to_gpu(A_host)
        A = torch.empty_like(A0)
        ctx = qr.context()
        stream = torch.cuda.current_stream()

        for _ in range(args.warmup):
            A.copy_(A0)
            qr.QR_2D(A, m, m, n, b, sync=False)
        ctx.sync()

        def run_steps(profiled):
            """K factorizations, each bracketed by events on the launch stream;
            with `profiled`, the library also brackets every kernel with events
            (its per-kernel durations feed the roofline and the breakdown)."""
            if profiled:
                ctx.profile(True)
                ctx.profile_reset()
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps)]
            torch.cuda.synchronize()
            w0 = time.perf_counter()
            for s in range(args.steps):
                A.copy_(A0)
                ev[s][0].record(stream)
                qr.QR_2D(A, m, m, n, b, sync=False)
                ev[s][1].record(stream)
            torch.cuda.synchronize()
            w = time.perf_counter() - w0
            ctx.sync()   # raises if any tree wait of these steps gave up
            pr = None
            if profiled:
                pr = ctx.profile_read()
                ctx.profile(False)
            return [a.elapsed_time(z) for a, z in ev], w, pr

        # the headline steps run without per-kernel events (they cost time of
        # their own); the same K steps are then repeated with them
        step_ms, wall, _ = run_steps(False)
        step_ms_prof, _, prof = run_steps(True)
        ms = float(np.mean(step_ms))
        F = qr_flops(m, n)
        gflops = F / (ms * 1e-3) / 1e9

        # dominant kernel class by time, its algorithmic flops/bytes per launch
        kinds = {k: v for k, v in prof.items() if v[3] > 0}
        dom = max(kinds, key=lambda k: kinds[k][0])
        dms, dfl, dby, dn = kinds[dom]
        is_mfma = dom.startswith("gemm")
        achieved = (dfl / (dms * 1e-3) / 1e12) if is_mfma else (dby / (dms * 1e-3) / 1e9)
        peak = PEAK_FP64_TFLOPS if is_mfma else PEAK_HBM_GBS
        # HBM bytes per launch from the committed counter runs (FETCH_SIZE and
        # WRITE_SIZE, calibrated on a 1 GiB copy) when they cover this kernel
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"qr2d_32768x4096_b128_{dom}_traffic.json")
        if os.path.exists(tf) and (m, n, b) == (32768, 4096, 128):
            with open(tf) as fh:
                tj = json.load(fh)
            if tj.get("kernel_class") == dom:
                traffic = tj["traffic_bytes_per_launch"]
        roofline = {
            "bound": "mfma" if is_mfma else "hbm",
            "kernel": dom,
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s" if is_mfma else "GB/s",
            "frac": round(achieved / peak, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(dby / dn),
            "avg_launch_ms": round(dms / dn, 5),
            "launches_per_step": dn / args.steps,
        }
        breakdown = {k: {"ms_per_step": round(v[0] / args.steps, 3), "launches_per_step": v[3] / args.steps,
                         "tflops": round(v[1] / (v[0] * 1e-3) / 1e12, 2) if v[0] > 0 else None}
                     for k, v in kinds.items()}

        out = {
            "metric": METRIC,
            "value": round(gflops, 2),
            "unit": "GFLOP/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": round(ms, 3),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "FP64 QR_2D (TSQR+HR panels, WY update) of a 32768x4096 N(0,1) matrix",
                       "m": m, "n": n, "b": b, "seed": args.seed},
            "qr_roofline_frac": round(gflops / 1e3 / PEAK_FP64_TFLOPS, 4),
            "roofline": roofline,
            "breakdown": breakdown,
            "wall_ms_per_step_incl_restore": round(wall / args.steps * 1e3, 3),
            "ms_per_step_with_kernel_events": round(float(np.mean(step_ms_prof)), 3),
        }

        if not args.no_accuracy:
            YR = qr.from_gpu(A, m, n)
            res, orth = accuracy(A_host, YR)
            out["accuracy"] = {"resid_over_n_eps": float(f"{res:.4g}"), "orth": float(f"{orth:.4g}"),
                               "probe": "randomized ‖Ax−QRx‖/(‖A‖‖x‖nε), ‖QᵀQv−v‖/‖v‖"}
        del A, A0
        torch.cuda.empty_cache()
    if not args.no_lsq:
        out["lsq_mixed"] = lsq_leg(qr, torch)
        torch.cuda.empty_cache()
    if not args.no_tree:
        out["tree_apply"] = tree_leg(qr, torch)
        torch.cuda.empty_cache()
    if not args.no_batch:
        out["batch"] = batch_leg(qr, torch, loop=not args.no_loop)
        torch.cuda.empty_cache()
    if not args.no_serve:
        out["serving"] = serve_leg(qr, torch)
        # the same stream with every request 64 columns wide: one shape, so
        # the server's batches are one width group each
        out["serving_n64"] = serve_leg(qr, torch, n_fixed=64)
        torch.cuda.empty_cache()
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
