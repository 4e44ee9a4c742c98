// Lean FP64 MFMA GEMM for the unmasked bulk of the compact-WY trailing
// update (alg/QR/qr_2d/qr_2d.cxx:259-275, the two dgemm calls of upd_A):
//   C = alpha·op(A)·B + beta·C   or   raw split-K partials of op(A)·B,
// with B k-contiguous (column-major K×N) and A either m-contiguous
// (column-major M×K, the update A −= Y·Z') or k-contiguous (Aᵀ with A
// column-major K×M, the product Z = Yᵀ·A).  No structural masks: the
// unit-lower top block of Y goes through the general kernel.
//
// 128×128 tile, 256 threads (4 waves of 64×64, 16 MFMA tiles each),
// 16-deep k slabs double-buffered in LDS with the next slab fetched by
// 16-byte loads into registers during the current one; a k-contiguous
// operand is kept k-contiguous in LDS so both the global loads and the LDS
// writes are 16 bytes wide.  Register use is bounded for two workgroups
// per CU, so one workgroup's loads and epilogue overlap the other's MFMAs.
// Tiles are ordered XCD by XCD, columns fastest, so the workgroups sharing
// a row block of A run back to back on the same L2.
#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

namespace {

typedef double d4l __attribute__((ext_vector_type(4)));
typedef double d2l __attribute__((ext_vector_type(2)));

constexpr int BM = 128, BN = 128, BK = 16;
constexpr int PM = BM + 16;  // pitch of an m-contiguous slab row [k][m]: rows k, k+1 fall on opposite bank halves
constexpr int PK = BK + 2;   // pitch of a k-contiguous slab row [idx][k]: 16 rows × 2 k cover all banks

struct LeanArgs {
  int64_t M, N, K;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  double alpha, beta;
  int vec;                   // operands allow 16-byte loads (even ld, aligned bases)
  int64_t kchunk;            // k per split (multiple of 16)
  int64_t split_stride;      // > 0: raw partial of split z at C + z·split_stride (ld M)
  int ntm, ntn, nsplit;      // tile grid and number of k splits
  int per_xcd;               // work items (split, tile) per XCD
};

}  // namespace

template <bool AK>
__global__ void __launch_bounds__(256, 2) lean_gemm_f64(LeanArgs g) {
  // LDS: A slab as [k][m] (AK = false) or [m][k] (AK = true); B as [n][k]
  __shared__ __align__(16) double As[2][AK ? BM * PK : BK * PM];
  __shared__ __align__(16) double Bs[2][BN * PK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // workgroups are dealt to the 8 XCDs round-robin; XCD x takes the
  // contiguous item range [x·per_xcd, (x+1)·per_xcd), items ordered split
  // by split with column tiles fastest, so an XCD's L2 serves one row block
  // (or one k chunk of Y) to many workgroups
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int tiles = g.ntm * g.ntn;
  const int u = xcd * g.per_xcd + slot;
  if (slot >= g.per_xcd || u >= tiles * g.nsplit) return;
  const int z = u / tiles, t = u - z * tiles;
  const int tm = t / g.ntn, tn = t - tm * g.ntn;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kbeg = (int64_t)z * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
  const bool full = g.vec && (m0 + BM <= g.M) && (n0 + BN <= g.N) && ((kend - kbeg) % BK == 0);
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;

  // per-thread slab elements (4 pairs of doubles per operand):
  //   m-contiguous A: k = kq + 4s, m = mp, mp+1 (a wave reads 1 KB of one column)
  //   k-contiguous  : idx = ri, k = kh + 2s, +1 (a thread reads 64 contiguous bytes of one column)
  const int kq = tid >> 6, mp = (tid & 63) * 2;
  const int ri = tid >> 1, kh = (tid & 1) * 8;
  d2l ra[4], rb[4];

  auto load = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (!AK) {
        const int64_t k = k0 + kq + 4 * s, m = m0 + mp;
        if (full) {
          ra[s] = *reinterpret_cast<const d2l*>(g.A + m + k * g.lda);
        } else {
          ra[s].x = (k < kend && m < g.M) ? g.A[m + k * g.lda] : 0.0;
          ra[s].y = (k < kend && m + 1 < g.M) ? g.A[m + 1 + k * g.lda] : 0.0;
        }
      } else {
        const int64_t k = k0 + kh + 2 * s, m = m0 + ri;
        if (full) {
          ra[s] = *reinterpret_cast<const d2l*>(g.A + k + m * g.lda);
        } else {
          ra[s].x = (k < kend && m < g.M) ? g.A[k + m * g.lda] : 0.0;
          ra[s].y = (k + 1 < kend && m < g.M) ? g.A[k + 1 + m * g.lda] : 0.0;
        }
      }
      const int64_t k = k0 + kh + 2 * s, n = n0 + ri;
      if (full) {
        rb[s] = *reinterpret_cast<const d2l*>(g.B + k + n * g.ldb);
      } else {
        rb[s].x = (k < kend && n < g.N) ? g.B[k + n * g.ldb] : 0.0;
        rb[s].y = (k + 1 < kend && n < g.N) ? g.B[k + 1 + n * g.ldb] : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (!AK) *reinterpret_cast<d2l*>(&As[buf][(kq + 4 * s) * PM + mp]) = ra[s];
      else *reinterpret_cast<d2l*>(&As[buf][ri * PK + kh + 2 * s]) = ra[s];
      *reinterpret_cast<d2l*>(&Bs[buf][ri * PK + kh + 2 * s]) = rb[s];
    }
  };

  d4l acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4l{0.0, 0.0, 0.0, 0.0};

  const int fr = lane >> 4, fc = lane & 15;
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        af[i] = AK ? As[buf][(wm0 + 16 * i + fc) * PK + kk + fr] : As[buf][(kk + fr) * PM + wm0 + 16 * i + fc];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][(wn0 + 16 * j + fc) * PK + kk + fr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
  };

  // lane l, register v of acc[i][j] holds C[m0+wm0+16i+(l&15)][n0+wn0+16j+(l>>4)+4v]
  const bool split = g.split_stride > 0;
  double* C = split ? g.C + (int64_t)z * g.split_stride : g.C;
  const int64_t ldc = split ? g.M : g.ldc;
  const bool has_beta = !split && g.beta != 0.0;
  const double alpha = split ? 1.0 : g.alpha;
  double cb[2][16];
  auto cload = [&](int i, double* dst) {
    const int64_t gm = m0 + wm0 + 16 * i + fc;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + fr + 4 * v;
        dst[4 * j + v] = (has_beta && gm < g.M && gn < g.N) ? C[gm + gn * ldc] : 0.0;
      }
  };

  const int64_t nslab = (kend - kbeg + BK - 1) / BK;
  if (nslab > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  for (int64_t sl = 0; sl + 1 < nslab; ++sl) {
    const int buf = (int)(sl & 1);
    load(kbeg + (sl + 1) * BK);
    compute(buf);
    store(buf ^ 1);
    __syncthreads();
  }
  // last slab: the first batch of C rows is fetched under its MFMAs
  cload(0, cb[0]);
  if (nslab > 0) compute((int)((nslab - 1) & 1));

  // epilogue: batch i+1 of C is in flight while batch i is combined and stored
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i + 1 < 4) cload(i + 1, cb[(i + 1) & 1]);
    const int64_t gm = m0 + wm0 + 16 * i + fc;
    if (gm >= g.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + fr + 4 * v;
        if (gn >= g.N) continue;
        double val = alpha * acc[i][j][v];
        if (has_beta) val += g.beta * cb[i & 1][4 * j + v];
        C[gm + gn * ldc] = val;
      }
  }
}

// C (M×N) = alpha·op(A)·B + beta·C over k ∈ [0, K); or, with scratch,
// raw partial sums of at most nsplit k chunks written to scratch + z·M·N
// (ld M); *used receives the number of partials written.
int lean_gemm(cqr_ctx* c, int ak, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* scratch, int nsplit,
              int kind, int* used) {
  if (used) *used = 0;
  if (M <= 0 || N <= 0 || K <= 0) return CQR_OK;
  LeanArgs g{};
  g.M = M; g.N = N; g.K = K; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
  g.alpha = alpha; g.beta = beta;
  g.vec = (lda % 2 == 0) && (ldb % 2 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0);
  g.ntm = (int)((M + BM - 1) / BM);
  g.ntn = (int)((N + BN - 1) / BN);
  const int64_t tiles = (int64_t)g.ntm * g.ntn;
  if (scratch) {
    nsplit = std::max(nsplit, 1);
    int64_t kc = (K + nsplit - 1) / nsplit;
    kc = (kc + BK - 1) / BK * BK;
    g.kchunk = kc;
    g.C = scratch;
    g.ldc = M;
    g.split_stride = M * N;
    nsplit = (int)((K + kc - 1) / kc);
  } else {
    g.kchunk = K;
    g.C = C;
    g.ldc = ldc;
    g.split_stride = 0;
    nsplit = 1;
  }
  if (used) *used = nsplit;
  g.nsplit = nsplit;
  if (tiles * nsplit >= (int64_t)1 << 30) return CQR_ERR_INVALID;
  g.per_xcd = (int)((tiles * nsplit + 7) / 8);
  const dim3 grid((unsigned)(8 * g.per_xcd));
  hipEvent_t ev = prof_begin(c);
  if (ak) hipLaunchKernelGGL(lean_gemm_f64<true>, grid, dim3(256), 0, c->stream, g);
  else hipLaunchKernelGGL(lean_gemm_f64<false>, grid, dim3(256), 0, c->stream, g);
  prof_end(c, ev, kind, 2.0 * M * N * K,
           8.0 * ((double)M * K + (double)K * N + (scratch ? nsplit : (beta != 0.0 ? 2.0 : 1.0)) * (double)M * N));
  return check_launch(c, "lean_gemm_f64");
}

}  // namespace cqr
