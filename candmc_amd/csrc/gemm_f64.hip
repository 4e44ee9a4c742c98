// FP64 MFMA GEMMs for the compact-WY trailing update on gfx950
// (alg/QR/qr_2d/qr_2d.cxx:224-282, upd_A):
//   Z  = Yᵀ·A          split-K over the rows (the column Allreduce of :265
//                      becomes a fixed-order sum of partial tiles)
//   Z' = Tᵀ·Z          (the dtrsm of :271 with T⁻ᵀ, done as a product with Tᵀ)
//   A −= Y·Z'          (:275)
// One kernel template serves all three: 128×128 block tile, 4 waves of
// 64×64, v_mfma_f64_16x16x4f64, LDS double buffering with a register
// prefetch of the next K slab.
#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

typedef double d4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
  int64_t m, n, k;
  double alpha, beta;
  const double* A;
  int64_t lda;
  int smA;
  const double* B;
  int64_t ldb;
  int smB;
  double* C;
  int64_t ldc;
  int64_t kchunk;        // rows of K per split (grid.z splits)
  int64_t split_stride;  // > 0: write raw partial sums to C + z·split_stride
};

__device__ __forceinline__ double gmask(double v, int64_t r, int64_t c, int sm) {
  switch (sm) {
    case SM_UNIT_LOWER: return r < c ? 0.0 : (r == c ? 1.0 : v);
    case SM_UPPER: return r > c ? 0.0 : v;
    case SM_LOWER: return r < c ? 0.0 : v;
    case SM_UNIT_UPPER: return r > c ? 0.0 : (r == c ? 1.0 : v);
    default: return v;
  }
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) mfma_gemm_f64(GemmArgs g) {
  constexpr int BM = 128, BN = 128, BK = 16, PAD = 4;
  constexpr int LDS_A = BM + PAD, LDS_B = BN + PAD;
  __shared__ double As[2][BK][LDS_A];
  __shared__ double Bs[2][BK][LDS_B];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = min(g.k, kbeg + g.kchunk);
  const double* __restrict__ A = g.A;
  const double* __restrict__ B = g.B;

  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[8], rb[8];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + 256 * r;
      int mm, kk;
      if (!TA) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < g.m && gk < kend) {
        const int64_t sr = TA ? gk : gm, sc = TA ? gm : gk;
        v = gmask(A[sr + sc * g.lda], sr, sc, g.smA);
      }
      ra[r] = v;
      int nn;
      if (!TB) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
      const int64_t gn = n0 + nn, gk2 = k0 + kk;
      double u = 0.0;
      if (gn < g.n && gk2 < kend) {
        const int64_t sr = TB ? gn : gk2, sc = TB ? gk2 : gn;
        u = gmask(B[sr + sc * g.ldb], sr, sc, g.smB);
      }
      rb[r] = u;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + 256 * r;
      int mm, kk;
      if (!TA) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
      As[buf][kk][mm] = ra[r];
      int nn;
      if (!TB) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
      Bs[buf][kk][nn] = rb[r];
    }
  };

  const int64_t ntiles = (kend - kbeg + BK - 1) / BK;
  if (ntiles > 0) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  const int fr = lane >> 4, fc = lane & 15;
  for (int64_t t = 0; t < ntiles; ++t) {
    const int buf = (int)(t & 1);
    if (t + 1 < ntiles) load(kbeg + (t + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[buf][kk + fr][wm0 + i * 16 + fc];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][kk + fr][wn0 + j * 16 + fc];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < ntiles) store(buf ^ 1);
    __syncthreads();
  }

  // lane l, register v holds C[m0+wm0+16i+(l&15)][n0+wn0+16j+(l>>4)+4v]
  double* C = g.C;
  const bool split = g.split_stride > 0;
  if (split) C += (int64_t)blockIdx.z * g.split_stride;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + wm0 + 16 * i + fc;
    if (gm >= g.m) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + fr + 4 * v;
        if (gn >= g.n) continue;
        double* cp = C + gm + gn * g.ldc;
        double val = acc[i][j][v];
        if (!split) {
          val *= g.alpha;
          if (g.beta != 0.0) val += g.beta * *cp;
        }
        *cp = val;
      }
    }
  }
}

// Z := Σ_z P_z in a fixed order (deterministic split-K reduction).
__global__ void splitk_reduce_kernel(const double* __restrict__ P, int64_t stride, int nsplit,
                                     int64_t m, int64_t n, double* __restrict__ Z, int64_t ldz,
                                     double alpha, double beta) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % m, cc = e / m;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += P[z * stride + r + cc * m];
    double* zp = Z + r + cc * ldz;
    double v = alpha * s;
    if (beta != 0.0) v += beta * *zp;
    *zp = v;
  }
}

static int launch_mfma(cqr_ctx* c, int ta, int tb, const GemmArgs& g, int nsplit) {
  dim3 grid((unsigned)((g.m + 127) / 128), (unsigned)((g.n + 127) / 128), (unsigned)nsplit);
  if (!ta && !tb) hipLaunchKernelGGL((mfma_gemm_f64<false, false>), grid, dim3(256), 0, c->stream, g);
  else if (ta && !tb) hipLaunchKernelGGL((mfma_gemm_f64<true, false>), grid, dim3(256), 0, c->stream, g);
  else if (!ta && tb) hipLaunchKernelGGL((mfma_gemm_f64<false, true>), grid, dim3(256), 0, c->stream, g);
  else hipLaunchKernelGGL((mfma_gemm_f64<true, true>), grid, dim3(256), 0, c->stream, g);
  return check_launch(c, "mfma_gemm_f64");
}

// Split-K choice: enough blocks to fill 256 CUs twice, chunks of ≥ 256 rows.
static int choose_splits(int64_t m, int64_t n, int64_t k, int64_t* kchunk) {
  const int64_t tiles = ((m + 127) / 128) * ((n + 127) / 128);
  int64_t s = 1;
  if (tiles < 512 && k > 512) {
    s = (512 + tiles - 1) / tiles;
    s = std::min<int64_t>(s, (k + 255) / 256);
    s = std::max<int64_t>(s, 1);
  }
  int64_t kc = (k + s - 1) / s;
  kc = (kc + 15) / 16 * 16;
  s = (k + kc - 1) / kc;
  *kchunk = kc;
  return (int)s;
}

// Generic C := alpha·op(A)·op(B) + beta·C; split-K through the workspace
// when the output is small and K long.  `scratch` must hold nsplit·m·n
// doubles when splitting (callers size it with gemm_f64_scratch).
static int gemm_f64_impl(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t lda, int smA, const double* B, int64_t ldb, int smB,
                         double beta, double* C, int64_t ldc, double* scratch) {
  if (m <= 0 || n <= 0) return CQR_OK;
  GemmArgs g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.smA = smA; g.B = B; g.ldb = ldb; g.smB = smB;
  int64_t kc;
  const int ns = scratch ? choose_splits(m, n, k, &kc) : 1;
  if (ns <= 1) {
    g.C = C; g.ldc = ldc; g.kchunk = std::max<int64_t>(k, 1); g.split_stride = 0;
    return launch_mfma(c, ta, tb, g, 1);
  }
  g.C = scratch; g.ldc = m; g.kchunk = kc; g.split_stride = m * n;
  CQR_TRY(launch_mfma(c, ta, tb, g, ns));
  const int gx = (int)std::min<int64_t>(2048, (m * n + 255) / 256);
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(gx), dim3(256), 0, c->stream, scratch, m * n, ns, m, n,
                     C, ldc, alpha, beta);
  return check_launch(c, "splitk_reduce");
}

size_t gemm_f64_scratch_doubles(int64_t m, int64_t n, int64_t k) {
  int64_t kc;
  const int ns = choose_splits(m, n, k, &kc);
  return ns > 1 ? (size_t)ns * m * n : 0;
}

int gemm_f64_ws(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double* A, int64_t lda, int smA, const double* B, int64_t ldb, int smB,
                double beta, double* C, int64_t ldc, double* scratch) {
  return gemm_f64_impl(c, ta, tb, m, n, k, alpha, A, lda, smA, B, ldb, smB, beta, C, ldc, scratch);
}

int gemm_f64(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
             const double* A, int64_t lda, int smA, const double* B, int64_t ldb, double beta,
             double* C, int64_t ldc) {
  const size_t sd = gemm_f64_scratch_doubles(m, n, k);
  double* scratch = nullptr;
  if (sd) {
    CQR_TRY(ensure_ws(c, sd * sizeof(double)));
    scratch = static_cast<double*>(c->ws);
  }
  return gemm_f64_impl(c, ta, tb, m, n, k, alpha, A, lda, smA, B, ldb, SM_NONE, beta, C, ldc, scratch);
}

// A (m×n) −= Y·Tt·(Yᵀ·A).  Workspace layout (doubles): partials | Z | Z'.
size_t wy_update_ws_doubles(int64_t m, int64_t n, int b) {
  return gemm_f64_scratch_doubles(b, n, m) + 2 * (size_t)b * n + 64;
}

int wy_update_f64_ws(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt,
                     double* A, int64_t lda, int64_t m, int64_t n, int b, double* ws) {
  if (m <= 0 || n <= 0 || b <= 0) return CQR_OK;
  const int smY = y_implicit ? SM_UNIT_LOWER : SM_NONE;
  double* part = ws;
  const size_t np = gemm_f64_scratch_doubles(b, n, m);
  double* Z = ws + ((np + 31) / 32) * 32;
  double* Zp = Z + ((size_t)b * n + 31) / 32 * 32;
  // Z = Yᵀ·A  (b×n)
  CQR_TRY(gemm_f64_impl(c, 1, 0, b, n, m, 1.0, Y, ldy, smY, A, lda, SM_NONE, 0.0, Z, b,
                        np ? part : nullptr));
  // Z' = Tᵀ·Z
  CQR_TRY(gemm_f64_impl(c, 0, 0, b, n, b, 1.0, Tt, b, SM_LOWER, Z, b, SM_NONE, 0.0, Zp, b, nullptr));
  // A −= Y·Z'
  return gemm_f64_impl(c, 0, 0, m, n, b, -1.0, Y, ldy, smY, Zp, b, SM_NONE, 1.0, A, lda, nullptr);
}

int wy_update_f64(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt,
                  double* A, int64_t lda, int64_t m, int64_t n, int b) {
  const size_t need = wy_update_ws_doubles(m, n, b);
  CQR_TRY(ensure_ws(c, need * sizeof(double)));
  return wy_update_f64_ws(c, Y, ldy, y_implicit, Tt, A, lda, m, n, b, static_cast<double*>(c->ws));
}

}  // namespace cqr
