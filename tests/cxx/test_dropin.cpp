// Drop-in check in the style of the reference's test/QR/test_hh_recon.cxx
// and test/QR/test_qr_2d.cxx: the same per-element drand48 inputs, calls
// through the reference signatures (host arrays), and the reference's
// acceptance metric Σ_i ‖(A − QR)(:,i)‖/‖A(:,i)‖ ≤ 1e-9 computed on the
// host with plain loops.  Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "candmc/CANDMC_QR.h"
#include "candmc_qr.h"

static double colrel_residual(const std::vector<double>& A, const std::vector<double>& YR, int64_t m,
                              int64_t n, bool flip_last) {
  std::vector<double> QR((size_t)m * n, 0.0);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i <= j && i < m; ++i) QR[i + j * m] = YR[i + j * m];
  if (flip_last) QR[(m - 1) + (n - 1) * m] *= -1.0;  // test/QR/test_qr_2d.cxx:135-137
  for (int64_t i = std::min(m, n) - 1; i >= 0; --i) {  // Q = H_1 ... H_n
    double nrm = 0.0;
    for (int64_t r = i + 1; r < m; ++r) nrm += YR[r + i * m] * YR[r + i * m];
    double tau = 2.0 / (nrm + 1.0);
    if (tau == 2.0) tau = 0.0;                          // alg/shared/util.cxx:451-456
    for (int64_t j = 0; j < n; ++j) {
      double w = QR[i + j * m];
      for (int64_t r = i + 1; r < m; ++r) w += YR[r + i * m] * QR[r + j * m];
      QR[i + j * m] -= tau * w;
      for (int64_t r = i + 1; r < m; ++r) QR[r + j * m] -= tau * w * YR[r + i * m];
    }
  }
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    double d = 0.0, a = 0.0;
    for (int64_t r = 0; r < m; ++r) {
      d += (A[r + j * m] - QR[r + j * m]) * (A[r + j * m] - QR[r + j * m]);
      a += A[r + j * m] * A[r + j * m];
    }
    s += std::sqrt(d / a);
  }
  return s;
}

static double seeded(int64_t row, int64_t col, int64_t m) {
  srand48(99900 + (row + col * m) * 61);
  return drand48();
}

int main() {
  int fails = 0;
  CommData_t cdt{nullptr, 1, 0, 0, 1};
  {  // test_hh_recon.cxx defaults at 4 ranks: b = 17, m = 34·4
    const int64_t P = 4, b = 17, m = 34 * P;
    std::vector<double> A((size_t)m * b), YR, W((size_t)b * b);
    for (int64_t c = 0; c < b; ++c)
      for (int64_t r = 0; r < m; ++r) A[r + c * m] = seeded(r, c, m);
    YR = A;
    hh_recon_qr(YR.data(), m, m, b, W.data(), 0, P, 0, 0, cdt);
    const double norm = colrel_residual(A, YR, m, b, false);
    printf("hh_recon_qr m=%lld b=%lld: ||A-QR||_2 = %.2E\n", (long long)m, (long long)b, norm);
    if (!(norm <= 1e-9)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test successful.\n");
  }
  {  // test_qr_2d.cxx input (drand48 − 0.5) through QR_2D and QR_2D_2D
    const int64_t m = 96, k = 48;
    std::vector<double> A((size_t)m * k);
    for (int64_t c = 0; c < k; ++c)
      for (int64_t r = 0; r < m; ++r) A[r + c * m] = seeded(r, c, m) - 0.5;
    pview pv{};
    pv.crow = cdt; pv.ccol = cdt; pv.cworld = cdt;
    std::vector<double> YR = A;
    QR_2D(YR.data(), m, m, k, 8, &pv, nullptr, 0);
    double norm = colrel_residual(A, YR, m, k, m == k);
    printf("QR_2D m=%lld k=%lld b=8: ||A-QR||_2 = %.2E\n", (long long)m, (long long)k, norm);
    if (!(norm <= 1e-9)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test successful.\n");
    YR = A;
    QR_2D_2D(YR.data(), m, m, k, 16, 8, &pv, nullptr, 0);
    norm = colrel_residual(A, YR, m, k, m == k);
    printf("QR_2D_2D m=%lld k=%lld b2=16 b=8: ||A-QR||_2 = %.2E\n", (long long)m, (long long)k, norm);
    if (!(norm <= 1e-9)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test successful.\n");
  }
  {  // square: the reference flips the last R entry before comparing
    const int64_t m = 64, k = 64;
    std::vector<double> A((size_t)m * k);
    srand48(7);
    for (auto& x : A) x = drand48() - 0.5;
    pview pv{};
    pv.crow = cdt; pv.ccol = cdt; pv.cworld = cdt;
    std::vector<double> YR = A;
    QR_2D_pipe(YR.data(), m, m, k, 16, &pv);
    const double norm = colrel_residual(A, YR, m, k, true);
    printf("QR_2D_pipe m=k=64 b=16: ||A-QR||_2 = %.2E\n", norm);
    if (!(norm <= 1e-9)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test successful.\n");
  }
  {  // test_qr_tree_2d.cxx on one rank: QR of [A, A] gives [Y\R_1, R_2] with R_2 = R_1
    const int64_t m = 96, k = 40, b = 8;
    std::vector<double> A((size_t)m * 2 * k);
    for (int64_t c = 0; c < k; ++c)
      for (int64_t r = 0; r < m; ++r) A[r + c * m] = A[r + (k + c) * m] = seeded(r, c, m) - 0.5;
    QR_tree_2D(A.data(), m, m, 2 * k, b, 0, 1, 0, 0, cdt, cdt, cdt, std::min(m, k));
    double f = 0.0;
    for (int64_t c = 0; c < k; ++c)
      for (int64_t r = 0; r < m; ++r) {
        const double l = A[r + c * m], x = A[r + (k + c) * m];
        if (r > c) f += x * x;
        else f += std::min((l - x) * (l - x), ((l - x) / l) * ((l - x) / l));
      }
    f = std::sqrt(f);
    printf("QR of [A,A] gives [Y/R_1,0/R_2], checking norm ||0/R_2-R_1||_F=%E\n", f);
    if (!(f < 1e-6)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test passed.\n");
  }
  {  // bitree_tsqr over 4 virtual ranks, then apply_tsqr_QT of the input: [R; 0]
    // (reference-sized tau and tree_data: b and 2b²(ln P + 2) doubles)
    const int64_t m = 256, b = 16, P = 4;
    std::vector<double> A((size_t)m * b), Y, R((size_t)b * b), tau((size_t)b);
    srand48(11);
    for (auto& x : A) x = drand48() - 0.5;
    std::vector<double> td((size_t)(2.0 * b * b * (std::log((double)P) + 2)));
    Y = A;
    bitree_tsqr(Y.data(), m, R.data(), tau.data(), m, b, 0, P, 0, 0, cdt, 1, td.data());
    std::vector<double> B = A;
    apply_tsqr_QT(Y.data(), m, tau.data(), B.data(), m, m, b, b, 0, P, 0, cdt, td.data());
    double d = 0.0, rmax = 0.0;
    for (int64_t c = 0; c < b; ++c)
      for (int64_t r = 0; r < m; ++r) {
        const double want = r <= c ? R[r + c * b] : 0.0;
        d = std::max(d, std::fabs(B[r + c * m] - want));
        rmax = std::max(rmax, std::fabs(want));
      }
    printf("apply_tsqr_QT of the factored panel: max |Q^T A - [R;0]| / max|R| = %.2E\n", d / rmax);
    if (!(d / rmax < 1e-12)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test passed.\n");
  }
  // test_bitree_tsqr.cxx:319 calls bitree_tsqr(A, mb, R, tree_tau, m, b, rank, P, root, req, cdt, 1):
  // no tree_data, each rank's block gets its dgeqrf [V\R] and the senders' top
  // triangles the node reflectors.  Rebuild Q·[R; 0] from A's blocks alone.
  for (int64_t P : {1, 4, 3}) {
    const int64_t b = 64, m = 8192;
    std::vector<double> A((size_t)m * b), Y, R((size_t)b * b + 1, 0.0), tau((size_t)b + 1);
    srand48(21 + P);
    for (auto& x : A) x = drand48() - 0.5;
    R[(size_t)b * b] = tau[b] = 12345.0;  // canaries
    Y = A;
    bitree_tsqr(Y.data(), m, R.data(), tau.data(), m, b, 0, P, 0, 0, cdt, 1);
    const int64_t nw = P * b > m ? m / b : P;
    std::vector<int64_t> off(nw + 1, 0);
    for (int64_t r = 0; r < nw; ++r) off[r + 1] = off[r] + ((m / b) / P) * b + (((m / b) % P) > r ? b : 0);
    off[nw] = m;
    std::vector<double> X((size_t)m * b, 0.0);
    for (int64_t c = 0; c < b; ++c)
      for (int64_t r = 0; r <= c; ++r) X[r + c * m] = R[r + c * b];
    // one Householder reflector u (rows rows[], values val[]) applied to X
    auto reflect = [&](const std::vector<int64_t>& rows, const std::vector<double>& val) {
      double nrm = 0.0;
      for (size_t q = 1; q < val.size(); ++q) nrm += val[q] * val[q];
      double t = 2.0 / (1.0 + nrm);
      if (t == 2.0) t = 0.0;
      for (int64_t c = 0; c < b; ++c) {
        double w = 0.0;
        for (size_t q = 0; q < rows.size(); ++q) w += val[q] * X[rows[q] + c * m];
        for (size_t q = 0; q < rows.size(); ++q) X[rows[q] + c * m] -= t * w * val[q];
      }
    };
    std::vector<std::vector<std::pair<int64_t, int64_t>>> levels;
    for (int64_t n = nw; n > 1; n = (n + 1) / 2) {
      levels.emplace_back();
      for (int64_t z = 0; z < n / 2; ++z) levels.back().push_back({z, z + (n + 1) / 2});
    }
    for (auto lv = levels.rbegin(); lv != levels.rend(); ++lv)
      for (auto pr : *lv)
        for (int64_t i = b - 1; i >= 0; --i) {
          std::vector<int64_t> rows{off[pr.first] + i};
          std::vector<double> val{1.0};
          for (int64_t r = 0; r <= i; ++r) {
            rows.push_back(off[pr.second] + r);
            val.push_back(Y[off[pr.second] + r + i * m]);
          }
          reflect(rows, val);
        }
    for (int64_t rk = 0; rk < nw; ++rk)
      for (int64_t i = b - 1; i >= 0; --i) {
        std::vector<int64_t> rows{off[rk] + i};
        std::vector<double> val{1.0};
        for (int64_t r = off[rk] + i + 1; r < off[rk + 1]; ++r) {
          rows.push_back(r);
          val.push_back(Y[r + i * m]);
        }
        reflect(rows, val);
      }
    double d = 0.0, a = 0.0;
    for (size_t e = 0; e < X.size(); ++e) {
      d += (X[e] - A[e]) * (X[e] - A[e]);
      a += A[e] * A[e];
    }
    const bool top = std::fabs(R[0] - Y[0]) == 0.0;  // rank 0's top triangle holds R
    printf("bitree_tsqr no tree_data P=%lld: ||Q[R;0] - A||_F/||A||_F = %.2E\n", (long long)P, std::sqrt(d / a));
    if (!(std::sqrt(d / a) < 1e-12) || !top || R[(size_t)b * b] != 12345.0 || tau[b] != 12345.0) {
      printf("TEST FAILED!!!\n");
      ++fails;
    } else {
      printf("Test passed.\n");
    }
  }
  {  // reference-sized buffers with canaries at 8192x64, P = 4 (tree_data path)
    const int64_t b = 64, m = 8192, P = 4, k = b;
    const size_t ntd = (size_t)(2.0 * b * b * (std::log((double)P) + 2));
    std::vector<double> A((size_t)m * b), Y, R((size_t)b * b), tau((size_t)b + 4, 777.0), td(ntd + 4, 777.0),
        Q((size_t)m * k);
    srand48(5);
    for (auto& x : A) x = drand48() - 0.5;
    Y = A;
    bitree_tsqr(Y.data(), m, R.data(), tau.data(), m, b, 0, P, 0, 0, cdt, 1, td.data());
    construct_Q1(Y.data(), m, tau.data(), Q.data(), m, m, b, k, 0, P, 0, cdt, td.data());
    double orth = 0.0, res = 0.0, an = 0.0;
    for (int64_t i = 0; i < k; ++i)
      for (int64_t j = 0; j < k; ++j) {
        double s = 0.0;
        for (int64_t r = 0; r < m; ++r) s += Q[r + i * m] * Q[r + j * m];
        orth = std::max(orth, std::fabs(s - (i == j ? 1.0 : 0.0)));
      }
    for (int64_t c = 0; c < b; ++c)
      for (int64_t r = 0; r < m; ++r) {
        double s = 0.0;
        for (int64_t q = 0; q <= c; ++q) s += Q[r + q * m] * R[q + c * b];
        res += (s - A[r + c * m]) * (s - A[r + c * m]);
        an += A[r + c * m] * A[r + c * m];
      }
    bool canary = true;
    for (int i = 0; i < 4; ++i) canary = canary && tau[b + i] == 777.0 && td[ntd + i] == 777.0;
    printf("construct_Q1 P=4 with reference-sized tau/tree_data: |Q^TQ-I| = %.2E, ||QR-A||/||A|| = %.2E, canaries %s\n",
           orth, std::sqrt(res / an), canary ? "intact" : "OVERWRITTEN");
    if (!(orth < 1e-12 && std::sqrt(res / an) < 1e-12 && canary)) { printf("TEST FAILED!!!\n"); ++fails; }
    else printf("Test passed.\n");
  }
  {  // one rank (construct_tsqr_Q.cxx:87-102): k = m, the whole Q by dorgqr
    const int64_t b = 20, m = 300, k = m;
    std::vector<double> A((size_t)m * b), Y, R((size_t)b * b), tau((size_t)b), td(64), Q((size_t)m * k);
    srand48(9);
    for (auto& x : A) x = drand48() - 0.5;
    Y = A;
    bitree_tsqr(Y.data(), m, R.data(), tau.data(), m, b, 0, 1, 0, 0, cdt, 1, td.data());
    construct_Q1(Y.data(), m, tau.data(), Q.data(), m, m, b, k, 0, 1, 0, cdt, td.data());
    double orth = 0.0, res = 0.0;
    for (int64_t i = 0; i < k; ++i)
      for (int64_t j = 0; j < k; ++j) {
        double s = 0.0;
        for (int64_t r = 0; r < m; ++r) s += Q[r + i * m] * Q[j * m + r];
        orth = std::max(orth, std::fabs(s - (i == j ? 1.0 : 0.0)));
      }
    for (int64_t c = 0; c < b; ++c)
      for (int64_t r = 0; r < m; ++r) {
        double s = 0.0;
        for (int64_t q = 0; q <= c; ++q) s += Q[r + q * m] * R[q + c * b];
        res = std::max(res, std::fabs(s - A[r + c * m]));
      }
    printf("one-rank construct_Q1 k=m=%lld: |Q^TQ-I| = %.2E, max|QR-A| = %.2E\n", (long long)m, orth, res);
    if (!(orth < 1e-13 && res < 1e-13)) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test passed.\n");
  }
  {  // the two-Y compute_invT_from_Y (qr_2d.cxx:76-106): T = Y1^T Y2
    const int64_t mb = 333, b1 = 17, b2 = 40, ldt = 20;
    std::vector<double> Y1((size_t)mb * b1), Y2((size_t)mb * b2), T((size_t)ldt * b2, -1.0);
    srand48(3);
    for (auto& x : Y1) x = drand48() - 0.5;
    for (auto& x : Y2) x = drand48() - 0.5;
    pview pv{};
    pv.crow = cdt; pv.ccol = cdt; pv.cworld = cdt;
    compute_invT_from_Y(Y1.data(), mb, Y2.data(), mb, T.data(), ldt, mb, b1, b2, &pv);
    double e = 0.0;
    for (int64_t j = 0; j < b2; ++j)
      for (int64_t i = 0; i < b1; ++i) {
        double s = 0.0;
        for (int64_t r = 0; r < mb; ++r) s += Y1[r + i * mb] * Y2[r + j * mb];
        e = std::max(e, std::fabs(s - T[i + j * ldt]));
      }
    const bool pad = T[b1] == -1.0;  // rows b1..ldt-1 untouched
    printf("compute_invT_from_Y (two Y): max error %.2E\n", e);
    if (!(e < 1e-12) || !pad) { printf("TEST FAILED!!!\n"); ++fails; } else printf("Test passed.\n");
  }
  return fails ? 1 : 0;
}
