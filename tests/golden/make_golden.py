"""Writes the golden fixtures in this directory (run in a container with scipy).

The reference repository ships no golden vectors: its tests recompute
LAPACK at run time.  These fixtures freeze LAPACK's answers (scipy's
dgeqrf / dgeqrt / dtpqrt) on the reference tests' own default inputs, so
the oracle and the HIP path are checked against fixed numbers:

  hh_recon_136x17.npz   test/QR/test_hh_recon.cxx defaults at 4 ranks
                        (b=17, m=34·P, drand48 per-element seeds :78-85)
  qr2d_8x4.npz          test/QR/test_qr_2d.cxx defaults on one rank
                        (m=8, k=4, b=2, drand48()−0.5 :96-103)
  qr2d_96x48.npz        the same generator at m=96, k=48 (b=8 in tests);
                        this generator gives numerically singular matrices
                        (cond ~1e16), so only the residual metric and the
                        rows of R above the first tiny pivot are comparable
  qr2d_gauss_160x96.npz seeded N(0,1) 160×96 (well conditioned), elementwise
  tpqrt_b20.npz         one tree node [R1; R2] at b=20 (two dtpqrt blocks)
  c1_summary.npz        8192×64 drand48 matrix (C1 shape): R diagonal,
                        τ and column norms of R (outputs too big to store)

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.linalg import lapack as lp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.candmc_oracle import hh_recon_test_matrix, qr_2d_test_matrix  # noqa: E402


def geqrf_full(A):
    qr, tau, _, info = lp.dgeqrf(np.asfortranarray(A))
    assert info == 0
    n = min(A.shape)
    _, t, info = lp.dgeqrt(n, np.asfortranarray(A))
    assert info == 0
    return qr, tau, t


def main():
    A = hh_recon_test_matrix(136, 17)
    qr, tau, t = geqrf_full(A)
    np.savez(os.path.join(HERE, "hh_recon_136x17.npz"), A=A, YR=qr, tau=tau, T=t)

    for m, k, name in [(8, 4, "qr2d_8x4.npz"), (96, 48, "qr2d_96x48.npz")]:
        A = qr_2d_test_matrix(m, k)
        qr, tau, t = geqrf_full(A)
        np.savez(os.path.join(HERE, name), A=A, YR=qr, tau=tau, T=t)

    from oracle.candmc_oracle import gaussian
    A = gaussian(160, 96, 4242)
    qr, tau, t = geqrf_full(A)
    np.savez(os.path.join(HERE, "qr2d_gauss_160x96.npz"), A=A, YR=qr, tau=tau, T=t)

    rng = np.random.Generator(np.random.PCG64(20240607))
    b = 20
    R1 = np.triu(rng.standard_normal((b, b)))
    R2 = np.triu(rng.standard_normal((b, b)))
    a, bb, t, info = lp.dtpqrt(b, 16, np.asfortranarray(R1), np.asfortranarray(R2))
    assert info == 0
    np.savez(os.path.join(HERE, "tpqrt_b20.npz"), R1=R1, R2=R2, R=np.triu(a), V2=np.triu(bb), T=t)

    A = hh_recon_test_matrix(8192, 64)
    qr, tau, t = geqrf_full(A)
    R = np.triu(qr[:64, :64])
    np.savez(os.path.join(HERE, "c1_summary.npz"), Rdiag=np.diag(R).copy(), tau=tau,
             Rcolnorm=np.linalg.norm(R, axis=0), Ycolnorm=np.linalg.norm(np.tril(qr, -1), axis=0))


if __name__ == "__main__":
    main()
