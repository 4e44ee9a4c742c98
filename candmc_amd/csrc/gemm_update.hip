// FP64 MFMA kernel for the rank-b trailing update of the compact-WY QR,
//   C (M×N) −= P (M×K) · Z (K×N),   K = b ≤ 128 (or an aggregate of two),
// the second dgemm of upd_A (alg/QR/qr_2d/qr_2d.cxx:275) with P = Y·Tᵀ
// (or Y·T, or Q1 for Y = Q1·U⁻¹) formed once per panel.  P is column-major
// (m-contiguous), Z column-major (k-contiguous), C column-major.
//
// K is small, so a workgroup's C tile is read once at the start straight
// into the MFMA accumulators and written once at the end: no epilogue
// read-modify-write, no second C buffer.  P is negated as it is staged, so
// every MFMA accumulates −P·Z onto C.  The tile is (64·WM)×(64·WN) with WM·WN
// waves of 64×64 (16 v_mfma_f64_16x16x4f64 tiles each); K runs in 16-deep
// slabs double-buffered in LDS, the next slab fetched into registers with
// 16-byte loads during the current one.  Items are dealt to the XCDs in
// contiguous ranges, column tiles fastest, so the workgroups of one XCD
// share a row block of P in its L2.
#include "cqr_device.h"
#include "cqr_internal.h"

namespace cqr {

namespace {

typedef double d4u __attribute__((ext_vector_type(4)));
typedef double d2u __attribute__((ext_vector_type(2)));

struct UpdArgs {
  int64_t M, N, K;
  const double* P;
  int64_t ldp;
  const double* Z;
  int64_t ldz;
  double* C;
  int64_t ldc;
  double alpha;        // C := beta·C + alpha·P·Z (alpha = −1 for the update)
  int beta1;           // 1: C is read (beta = 1); 0: C is overwritten (beta = 0)
  // with beta = 0 only: columns [n1, N) go to C3 (column gn − n1, ld ldc3)
  // instead of C, and columns [0, n1) also to C2 (ld ldc2) when non-null
  int64_t n1;
  double* C2;
  int64_t ldc2;
  double* C3;
  int64_t ldc3;
  int vec;             // 16-byte loads allowed (even lds, aligned bases)
  int ntm, ntn;        // tile grid
  int per_xcd;         // items per XCD
};

constexpr int UBK = 16;

}  // namespace

// FULL: the tile lies inside M×N, K is a multiple of 16 and 16-byte loads
// are allowed, so every access is unconditional (straight-line code lets
// all of a slab's loads be in flight together).
template <int WM, int WN, bool FULL>
__device__ __forceinline__ void upd_tile(const UpdArgs& g, int64_t m0, int64_t n0, double* __restrict__ Psm,
                                         double* __restrict__ Zsm) {
  constexpr int NT = 64 * WM * WN;
  constexpr int BM = 64 * WM, BN = 64 * WN;
  constexpr int PMP = BM + 16;          // pitch of a P slab row [k][m]
  constexpr int PZ = UBK + 2;           // pitch of a Z slab row [n][k]
  constexpr int NPA = UBK * BM / 2 / NT;  // 16-byte P loads per thread per slab
  constexpr int NZB = UBK * BN / 2 / NT;  // 16-byte Z loads per thread per slab
  static_assert(NPA >= 1 && NZB >= 1, "tile too small for the thread count");
  double(*Ps)[UBK * PMP] = reinterpret_cast<double(*)[UBK * PMP]>(Psm);
  double(*Zs)[BN * PZ] = reinterpret_cast<double(*)[BN * PZ]>(Zsm);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr bool full = FULL;
  const int wm0 = (wave % WM) * 64, wn0 = (wave / WM) * 64;
  const int fr = lane >> 4, fc = lane & 15;

  // accumulators start as C: lane l, register v of acc[i][j] holds
  // C[m0+wm0+16i+(l&15)][n0+wn0+16j+(l>>4)+4v]
  d4u acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + wm0 + 16 * i + fc;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + fr + 4 * v;
        acc[i][j][v] = (g.beta1 && (full || (gm < g.M && gn < g.N))) ? g.C[gm + gn * g.ldc] : 0.0;
      }
  }

  // slab element ownership: P as pairs along m, Z as pairs along k
  d2u rp[NPA], rz[NZB];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < NPA; ++s) {
      const int e = tid + NT * s;              // pair index in the BM×16 slab
      const int kk = e / (BM / 2), mp = (e % (BM / 2)) * 2;
      const int64_t k = k0 + kk, m = m0 + mp;
      if (full) {
        rp[s] = *reinterpret_cast<const d2u*>(g.P + m + k * g.ldp);
      } else {
        rp[s].x = (k < g.K && m < g.M) ? g.P[m + k * g.ldp] : 0.0;
        rp[s].y = (k < g.K && m + 1 < g.M) ? g.P[m + 1 + k * g.ldp] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < NZB; ++s) {
      const int e = tid + NT * s;              // pair index in the 16×BN slab
      const int nn = e / (UBK / 2), kp = (e % (UBK / 2)) * 2;
      const int64_t k = k0 + kp, n = n0 + nn;
      if (full) {
        rz[s] = *reinterpret_cast<const d2u*>(g.Z + k + n * g.ldz);
      } else {
        rz[s].x = (k < g.K && n < g.N) ? g.Z[k + n * g.ldz] : 0.0;
        rz[s].y = (k + 1 < g.K && n < g.N) ? g.Z[k + 1 + n * g.ldz] : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int s = 0; s < NPA; ++s) {
      const int e = tid + NT * s;
      const int kk = e / (BM / 2), mp = (e % (BM / 2)) * 2;
      d2u x = rp[s];
      x.x *= g.alpha;
      x.y *= g.alpha;
      *reinterpret_cast<d2u*>(&Ps[buf][kk * PMP + mp]) = x;
    }
#pragma unroll
    for (int s = 0; s < NZB; ++s) {
      const int e = tid + NT * s;
      const int nn = e / (UBK / 2), kp = (e % (UBK / 2)) * 2;
      *reinterpret_cast<d2u*>(&Zs[buf][nn * PZ + kp]) = rz[s];
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < UBK; kk += 4) {
      double af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = Ps[buf][(kk + fr) * PMP + wm0 + 16 * i + fc];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Zs[buf][(wn0 + 16 * j + fc) * PZ + kk + fr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
  };

  const int nslab = (int)((g.K + UBK - 1) / UBK);
  load(0);
  store(0);
  __syncthreads();
  // the first slab on its own: its MFMAs wait once for the C loads, so the
  // loop below only ever waits for its own slab loads
  if (nslab > 1) load(UBK);
  compute(0);
  if (nslab > 1) store(1);
  __syncthreads();
  for (int sl = 1; sl < nslab; ++sl) {
    const int buf = sl & 1;
    if (sl + 1 < nslab) load((int64_t)(sl + 1) * UBK);
    compute(buf);
    if (sl + 1 < nslab) store(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + wm0 + 16 * i + fc;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + fr + 4 * v;
        if (full || (gm < g.M && gn < g.N)) {
          const double val = acc[i][j][v];
          if (gn < g.n1) {
            g.C[gm + gn * g.ldc] = val;
            if (g.C2) g.C2[gm + gn * g.ldc2] = val;
          } else {
            g.C3[gm + (gn - g.n1) * g.ldc3] = val;
          }
        }
      }
  }
}

template <int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN, (WM * WN <= 4) ? 2 : 1) upd_gemm_f64(UpdArgs g) {
  constexpr int BM = 64 * WM, BN = 64 * WN;
  __shared__ __align__(16) double Ps[2 * UBK * (BM + 16)];
  __shared__ __align__(16) double Zs[2 * BN * (UBK + 2)];
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int tiles = g.ntm * g.ntn;
  const int u = xcd * g.per_xcd + slot;
  if (slot >= g.per_xcd || u >= tiles) return;
  const int tm = u / g.ntn, tn = u - tm * g.ntn;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  if (g.vec && (m0 + BM <= g.M) && (n0 + BN <= g.N) && (g.K % UBK == 0)) upd_tile<WM, WN, true>(g, m0, n0, Ps, Zs);
  else upd_tile<WM, WN, false>(g, m0, n0, Ps, Zs);
}

static int launch_update(cqr_ctx* c, UpdArgs& g, int kind, int shape) {
  g.vec = (g.ldp % 2 == 0) && (g.ldz % 2 == 0) && ((uintptr_t)g.P % 16 == 0) && ((uintptr_t)g.Z % 16 == 0);
  const int bm = shape == 1 ? 256 : 128, bn = shape == 2 ? 256 : 128;
  g.ntm = (int)((g.M + bm - 1) / bm);
  g.ntn = (int)((g.N + bn - 1) / bn);
  const int64_t tiles = (int64_t)g.ntm * g.ntn;
  if (tiles >= (int64_t)1 << 30) return set_err(c, CQR_ERR_UNSUPPORTED, "update_gemm: too many tiles");
  g.per_xcd = (int)((tiles + 7) / 8);
  const dim3 grid((unsigned)(8 * g.per_xcd));
  hipEvent_t ev = prof_begin(c);
  if (shape == 1) hipLaunchKernelGGL((upd_gemm_f64<4, 2>), grid, dim3(512), 0, c->stream, g);
  else if (shape == 2) hipLaunchKernelGGL((upd_gemm_f64<2, 4>), grid, dim3(512), 0, c->stream, g);
  else hipLaunchKernelGGL((upd_gemm_f64<2, 2>), grid, dim3(256), 0, c->stream, g);
  const double outs = (g.beta1 ? 2.0 : 1.0) + (g.C2 ? (double)g.n1 / g.N : 0.0);
  prof_end(c, ev, kind, 2.0 * g.M * g.N * g.K, 8.0 * ((double)g.M * g.K + (double)g.K * g.N + outs * g.M * g.N));
  return check_launch(c, "upd_gemm_f64");
}

// C (M×N) := beta·C + alpha·P·Z with K ≤ 256 and beta ∈ {0, 1} (alpha = −1,
// beta = 1: the trailing update).  shape: 0 = 128×128 tiles (4 waves, two
// workgroups per CU), 1 = 256×128 (8 waves), 2 = 128×256 (8 waves).
int update_gemm(cqr_ctx* c, int64_t M, int64_t N, int64_t K, double alpha, const double* P, int64_t ldp,
                const double* Z, int64_t ldz, double beta, double* C, int64_t ldc, int kind, int shape) {
  if (M <= 0 || N <= 0) return CQR_OK;
  if (K > 256 || K < 1) return set_err(c, CQR_ERR_UNSUPPORTED, "update_gemm: K = %lld outside [1, 256]", (long long)K);
  if (beta != 0.0 && beta != 1.0) return set_err(c, CQR_ERR_UNSUPPORTED, "update_gemm: beta must be 0 or 1");
  UpdArgs g{};
  g.M = M; g.N = N; g.K = K; g.P = P; g.ldp = ldp; g.Z = Z; g.ldz = ldz; g.C = C; g.ldc = ldc; g.alpha = alpha;
  g.beta1 = beta != 0.0;
  g.n1 = N;
  return launch_update(c, g, kind, shape);
}

// P (M×K) times [Z1 | Z2] (K×N, ld ldz, Z2 = its columns [n1, N)) into
// three destinations at once: C1 := P·Z1 (ld ldc1) and, if C1b is
// non-null, a copy in C1b (ld ldc1b); C2 := P·Z2 (ld ldc2).  One pass over
// P for two products that share it (Y = Q1·U⁻¹ with the explicit Y, and
// Y·Tᵀ = Q1·(U⁻¹·Tᵀ)).
int update_gemm_split(cqr_ctx* c, int64_t M, int64_t N, int64_t n1, int64_t K, const double* P, int64_t ldp,
                      const double* Z, int64_t ldz, double* C1, int64_t ldc1, double* C1b, int64_t ldc1b, double* C2,
                      int64_t ldc2, int kind) {
  if (M <= 0 || N <= 0) return CQR_OK;
  if (K > 256 || K < 1 || n1 < 0 || n1 > N)
    return set_err(c, CQR_ERR_UNSUPPORTED, "update_gemm_split: K = %lld, n1 = %lld", (long long)K, (long long)n1);
  UpdArgs g{};
  g.M = M; g.N = N; g.K = K; g.P = P; g.ldp = ldp; g.Z = Z; g.ldz = ldz; g.C = C1; g.ldc = ldc1; g.alpha = 1.0;
  g.beta1 = 0;
  g.n1 = n1; g.C2 = C1b; g.ldc2 = ldc1b; g.C3 = C2; g.ldc3 = ldc2;
  return launch_update(c, g, kind, 0);
}

// ---------------------------------------------------------------------------
// Batched Y = Q1·U⁻¹ below the top b rows of every problem of a group
// (hh_recon.cxx:294-301 per problem): the same tile with per-block problem
// lookup.  A block's tile is a 128-row block of one problem times all of
// its b ≤ 128 columns (one tile column: 128 wide for b > 64, else 64), with
// K = b; the blocks of a problem are consecutive, so they share its U⁻¹
// in one XCD's L2.  Block u → (problem, row tile) through the segment
// table (one segment per problem with m > b).
// ---------------------------------------------------------------------------
namespace {
struct YBatchArgs {
  const BatchProb* bp;
  const BatchSeg* seg;
  int nseg, b, total, per_xcd;
  const double* Y;      // the group's Q1 rows (problem p at Y + yoff, ld m)
  const double* Uinv;   // b×b per problem, strictly lower part zero
  double* A;            // the group's A (problem p at A + aoff, ld m)
};
}  // namespace

template <int WN>
__global__ void __launch_bounds__(128 * WN, WN == 2 ? 2 : 1) upd_y_batched_f64(YBatchArgs a) {
  constexpr int BM = 128, BN = 64 * WN;
  __shared__ __align__(16) double Ps[2 * UBK * (BM + 16)];
  __shared__ __align__(16) double Zs[2 * BN * (UBK + 2)];
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int u = xcd * a.per_xcd + slot;
  if (slot >= a.per_xcd || u >= a.total) return;
  const BatchSeg S = a.seg[seg_find(a.seg, a.nseg, u)];
  const BatchProb P = a.bp[S.prob];
  const int b = a.b;
  UpdArgs g{};
  g.M = P.m - b;
  g.N = b;
  g.K = b;
  g.P = a.Y + P.yoff + b;
  g.ldp = P.m;
  g.Z = a.Uinv + (int64_t)S.prob * b * b;
  g.ldz = b;
  g.C = a.A + P.aoff + b;
  g.ldc = P.m;
  g.alpha = 1.0;
  g.n1 = b;
  // 16-byte loads: even leading dimensions and offsets on 16-byte bases
  g.vec = (P.m % 2 == 0) && (b % 2 == 0) && ((P.yoff + b) % 2 == 0);
  const int64_t m0 = (int64_t)(S.local0 + (u - S.g0)) * BM;
  if (g.vec && m0 + BM <= g.M && BN <= b && b % UBK == 0) upd_tile<2, WN, true>(g, m0, 0, Ps, Zs);
  else upd_tile<2, WN, false>(g, m0, 0, Ps, Zs);
}

int batched_y_tile_rows() { return 128; }

int update_y_batched(cqr_ctx* c, const BatchProb* bp, const BatchSeg* seg, int nseg, int total_tiles, int b,
                     const double* Y, const double* Uinv, double* A) {
  if (total_tiles <= 0) return CQR_OK;
  if (b < 1 || b > 128) return set_err(c, CQR_ERR_UNSUPPORTED, "update_y_batched: b = %d", b);
  YBatchArgs a{bp, seg, nseg, b, total_tiles, (total_tiles + 7) / 8, Y, Uinv, A};
  const dim3 grid((unsigned)(8 * a.per_xcd));
  hipEvent_t ev = prof_begin(c);
  if (b > 64) hipLaunchKernelGGL((upd_y_batched_f64<2>), grid, dim3(256), 0, c->stream, a);
  else hipLaunchKernelGGL((upd_y_batched_f64<1>), grid, dim3(128), 0, c->stream, a);
  prof_end(c, ev, PK_GEMM_SMALL, 0.0, 0.0);
  return check_launch(c, "upd_y_batched_f64");
}

}  // namespace cqr
