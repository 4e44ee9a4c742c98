// FP64 MFMA GEMMs for the compact-WY trailing update on gfx950
// (alg/QR/qr_2d/qr_2d.cxx:224-282, upd_A):
//   Z  = Yᵀ·A          split-K over the rows (the column Allreduce of :265
//                      becomes a fixed-order sum of partial tiles)
//   Z' = Tᵀ·Z          (the dtrsm of :271 with T⁻ᵀ, done as a product with Tᵀ)
//   A −= Y·Z'          (:275)
// One kernel template serves all three: 128×128 block tile, 4 waves of
// 64×64, v_mfma_f64_16x16x4f64, LDS double buffering with a register
// prefetch of the next K slab.  The ROLE template argument only tags the
// instantiation (0 general, 1 Z = Yᵀ·A, 2 A −= Y·Z', 3 batched per-leaf
// products) so that kernel traces attribute time to each use.
#include "cqr_device.h"
#include "cqr_internal.h"

#include <type_traits>

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
  // batched form (grid.z = batch index, no split): per-z operand addresses
  int batched;
  int a_mode, b_mode, c_mode;  // BM_STRIDED / BM_LEAFROWS / BM_SHIFTED
  int64_t a_zs, b_zs, c_zs;
  int64_t pm;                  // panel rows for BM_LEAFROWS
  int pL, zoff;
  int m_leaf;                  // m = rows of leaf z
  int smC;                     // SM_UPPER: store zeros below the diagonal
  // batched by problem (BM_PROB_*): per-z addresses and sizes from a table
  const BatchProb* bp;
  int64_t prow;                // row shift inside the problem's block
  int m_prob;                  // m = rows of problem z − prow
};

__device__ __forceinline__ int64_t leaf_row0(int64_t z, int64_t m, int64_t L) { return (z * m) / L; }

__device__ __forceinline__ const double* zaddr(const double* p, int mode, int64_t zs, int z, const GemmArgs& g) {
  if (mode == BM_LEAFROWS) return p + leaf_row0(z, g.pm, g.pL);
  if (mode == BM_SHIFTED) return p + (int64_t)(z + g.zoff) * zs;
  if (mode == BM_PROB_Y) return p + g.bp[z].yoff + g.prow;
  if (mode == BM_PROB_A) return p + g.bp[z].aoff + g.prow;
  if (mode == BM_PROB_T) return p + g.bp[z].toff;
  return p + (int64_t)z * zs;
}
// leading dimension of a per-problem operand (its block is m×·, ld m)
__device__ __forceinline__ int64_t zld(int64_t ld, int mode, int z, const GemmArgs& g) {
  return (mode == BM_PROB_Y || mode == BM_PROB_A) ? g.bp[z].m : ld;
}

__device__ __forceinline__ double gmask(double v, int64_t r, int64_t c, int sm) {
  switch (sm) {
    case SM_UNIT_LOWER: return r < c ? 0.0 : (r == c ? 1.0 : v);
    case SM_UPPER: return r > c ? 0.0 : v;
    case SM_LOWER: return r < c ? 0.0 : v;
    case SM_UNIT_UPPER: return r > c ? 0.0 : (r == c ? 1.0 : v);
    default: return v;
  }
}

// Does a structural mask change any element of the stored block
// rows [r_lo, r_hi] × cols [c_lo, c_hi]?
__device__ __forceinline__ bool mask_active(int sm, int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi) {
  switch (sm) {
    case SM_UNIT_LOWER: return r_lo <= c_hi;
    case SM_LOWER: return r_lo < c_hi;
    case SM_UPPER: return r_hi > c_lo;
    case SM_UNIT_UPPER: return r_hi >= c_lo;
    default: return false;
  }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// PREC = CQR_PREC_F64: FP64 MFMA.  CQR_PREC_F32: the FP64 operands are
// rounded to FP32 while they are staged in LDS and multiplied on FP32 MFMA
// (v_mfma_f32_16x16x4f32) with FP32 accumulation; the result is added to
// the FP64 output in FP64.

__device__ __forceinline__ unsigned short to_bf16_bits(double x) {
  return __builtin_bit_cast(unsigned short, (__bf16)(float)x);
}

// CQR_PREC_BF16: operands rounded to BF16 into k-contiguous LDS rows (a
// lane's eight consecutive k values are one 16-byte read), 32-deep slabs
// multiplied on gfx950's v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

template <bool TA, bool TB, int ROLE, int PREC = CQR_PREC_F64>
__global__ void __launch_bounds__(256) mfma_gemm_f64(GemmArgs g) {
  constexpr bool LP = PREC != CQR_PREC_F64, BF = PREC == CQR_PREC_BF16;
  // a slab is NH halves of 16 k each; every thread stages 8 values of A and
  // 8 of B per half
  constexpr int NH = BF ? 2 : 1;
  constexpr int BM = 128, BN = 128, BK = 16 * NH, PAD = 4, KP = BK + 8;
  constexpr int LDS_A = BM + PAD, LDS_B = BN + PAD;
  typedef typename std::conditional<LP, float, double>::type ST;
  typedef typename std::conditional<LP, f4v, d4>::type AT;
  constexpr size_t SZ_STD = 2 * BK * (LDS_A + LDS_B) * sizeof(ST);
  constexpr size_t SZ_BF = 2 * (BM + BN) * KP * sizeof(unsigned short);
  __shared__ __align__(16) unsigned char smem[BF ? SZ_BF : SZ_STD];
  auto As = reinterpret_cast<ST (*)[BK][LDS_A]>(smem);
  auto Bs = reinterpret_cast<ST (*)[BK][LDS_B]>(smem + 2 * BK * LDS_A * sizeof(ST));
  auto As16 = reinterpret_cast<unsigned short (*)[BM][KP]>(smem);
  auto Bs16 = reinterpret_cast<unsigned short (*)[BN][KP]>(smem + 2 * BM * KP * sizeof(unsigned short));

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const double* __restrict__ A = g.A;
  const double* __restrict__ B = g.B;
  double* C = g.C;
  int64_t Mrows = g.m, kbeg, kend;
  int64_t lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  if (g.batched) {
    const int z = blockIdx.z;
    A = zaddr(g.A, g.a_mode, g.a_zs, z, g);
    B = zaddr(g.B, g.b_mode, g.b_zs, z, g);
    C = const_cast<double*>(zaddr(g.C, g.c_mode, g.c_zs, z, g));
    lda = zld(lda, g.a_mode, z, g);
    ldb = zld(ldb, g.b_mode, z, g);
    ldc = zld(ldc, g.c_mode, z, g);
    if (g.m_leaf) Mrows = leaf_row0(z + 1, g.pm, g.pL) - leaf_row0(z, g.pm, g.pL);
    if (g.m_prob) Mrows = g.bp[z].m - g.prow;
    kbeg = 0;
    kend = g.k;
    if (m0 >= Mrows) return;   // whole block out of this batch entry's rows
  } else {
    kbeg = (int64_t)blockIdx.z * g.kchunk;
    kend = min(g.k, kbeg + g.kchunk);
  }
  const bool full_mn = (m0 + BM <= Mrows) && (n0 + BN <= g.n);

  // Each thread loads 8 elements of the A slab and 8 of the B slab:
  //   K-contiguous operand: k = k0 + (tid & 15), index = base + (tid >> 4) + 16r
  //   index-contiguous operand: index = base + (tid & 127), k = k0 + (tid >> 7) + 2r
  const int ak = TA ? (tid & 15) : (tid >> 7);
  const int am = TA ? (tid >> 4) : (tid & 127);
  const int bk = TB ? (tid >> 7) : (tid & 15);
  const int bn = TB ? (tid & 127) : (tid >> 4);
  const int64_t a_rstep = TA ? 16 * lda : 2 * lda;   // between the 8 elements
  const int64_t b_rstep = TB ? 2 * ldb : 16 * ldb;
  const double* pa = TA ? A + (m0 + am) * lda + ak : A + (m0 + am) + (int64_t)ak * lda;
  const double* pb = TB ? B + (n0 + bn) + (int64_t)bk * ldb : B + bk + (n0 + bn) * ldb;
  const int64_t a_kstep = TA ? 1 : lda;                 // per unit of k
  const int64_t b_kstep = TB ? ldb : 1;

  AT acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = AT{0, 0, 0, 0};

  double ra[8 * NH], rb[8 * NH];
  auto load_half = [&](int64_t k0, int h) {
    const bool kfull = k0 + 16 <= kend;
    const bool mA = TA ? mask_active(g.smA, k0, k0 + 15, m0, m0 + BM - 1)
                       : mask_active(g.smA, m0, m0 + BM - 1, k0, k0 + 15);
    const bool mB = TB ? mask_active(g.smB, n0, n0 + BN - 1, k0, k0 + 15)
                       : mask_active(g.smB, k0, k0 + 15, n0, n0 + BN - 1);
    const double* qa = pa + k0 * a_kstep;
    const double* qb = pb + k0 * b_kstep;
    if (full_mn && kfull && !mA && !mB) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        ra[8 * h + r] = qa[r * a_rstep];
        rb[8 * h + r] = qb[r * b_rstep];
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int64_t gm = m0 + am + (TA ? 16 * r : 0), gk = k0 + ak + (TA ? 0 : 2 * r);
      double v = 0.0;
      if (gm < Mrows && gk < kend) {
        v = qa[r * a_rstep];
        if (mA) v = gmask(v, TA ? gk : gm, TA ? gm : gk, g.smA);
      }
      ra[8 * h + r] = v;
      const int64_t gn = n0 + bn + (TB ? 0 : 16 * r), gk2 = k0 + bk + (TB ? 2 * r : 0);
      double u = 0.0;
      if (gn < g.n && gk2 < kend) {
        u = qb[r * b_rstep];
        if (mB) u = gmask(u, TB ? gn : gk2, TB ? gk2 : gn, g.smB);
      }
      rb[8 * h + r] = u;
    }
  };
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int h = 0; h < NH; ++h) load_half(k0 + 16 * h, h);
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int h = 0; h < NH; ++h)
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if constexpr (BF) {
          As16[buf][am + (TA ? 16 * r : 0)][16 * h + ak + (TA ? 0 : 2 * r)] = to_bf16_bits(ra[8 * h + r]);
          Bs16[buf][bn + (TB ? 0 : 16 * r)][16 * h + bk + (TB ? 2 * r : 0)] = to_bf16_bits(rb[8 * h + r]);
        } else {
          As[buf][ak + (TA ? 0 : 2 * r)][am + (TA ? 16 * r : 0)] = (ST)ra[r];
          Bs[buf][bk + (TB ? 2 * r : 0)][bn + (TB ? 0 : 16 * r)] = (ST)rb[r];
        }
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
    if constexpr (BF) {
      // one K = 32 product per tile pair: lane l supplies k = 8(l>>4)..+7
      bf8v af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = *reinterpret_cast<const bf8v*>(&As16[buf][wm0 + i * 16 + fc][8 * fr]);
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = *reinterpret_cast<const bf8v*>(&Bs16[buf][wn0 + j * 16 + fc][8 * fr]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j], af[i], acc[i][j], 0, 0, 0);
    } else
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      ST af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[buf][kk + fr][wm0 + i * 16 + fc];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][kk + fr][wn0 + j * 16 + fc];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if constexpr (LP)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
        }
    }
    if (t + 1 < ntiles) store(buf ^ 1);
    __syncthreads();
  }

  // lane l, register v holds C[m0+wm0+16i+(l&15)][n0+wn0+16j+c(l,v)] with
  // c = (l>>4)+4v for the FP64 MFMA and c = 4(l>>4)+v for the FP32 one
  // (operands are passed swapped, so each lane's values share one row)
  const bool split = g.split_stride > 0;
  if (split) C += (int64_t)blockIdx.z * g.split_stride;
  const bool has_beta = !split && g.beta != 0.0;
  const double alpha = split ? 1.0 : g.alpha;
  // C is read (beta ≠ 0) sixteen values at a time, all loads of a group
  // issued before any of its stores, so a tile pays four memory latencies
  // rather than one per element.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + wm0 + 16 * i + fc;
    const bool row_ok = full_mn || gm < Mrows;
    double cold[16];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + (LP ? 4 * fr + v : fr + 4 * v);
        const bool ok = row_ok && (full_mn || gn < g.n);
        cold[4 * j + v] = (has_beta && ok) ? C[gm + gn * ldc] : 0.0;
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gn = n0 + wn0 + 16 * j + (LP ? 4 * fr + v : fr + 4 * v);
        if (!row_ok || !(full_mn || gn < g.n)) continue;
        double val = alpha * (double)acc[i][j][v];
        if (has_beta) val += g.beta * cold[4 * j + v];
        if (g.smC == SM_UPPER && gm > gn) val = 0.0;
        C[gm + gn * ldc] = val;
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

// Few outputs over many partial slots (the next panel's b×b-wide Yᵀ·A): 64
// outputs per workgroup, four groups of slots per output summed in a fixed
// order, so the grid covers the chip even for a single 128×128 tile.
__global__ void __launch_bounds__(256) splitk_reduce_grouped_kernel(const double* __restrict__ P, int64_t stride,
                                                                    int nsplit, int64_t m, int64_t n,
                                                                    double* __restrict__ Z, int64_t ldz,
                                                                    double alpha, double beta) {
  __shared__ double part[4][64];
  const int t = threadIdx.x, el = t & 63, g = t >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + el;
  const bool ok = e < m * n;
  double s = 0.0;
  if (ok)
    for (int z = g; z < nsplit; z += 4) s += P[z * stride + e];
  part[g][el] = s;
  __syncthreads();
  if (g == 0 && ok) {
    const double tot = (part[0][el] + part[1][el]) + (part[2][el] + part[3][el]);
    double* zp = Z + (e % m) + (e / m) * ldz;
    double v = alpha * tot;
    if (beta != 0.0) v += beta * *zp;
    *zp = v;
  }
}

// Z (m×n, ld ldz) = alpha·Σ_z P_z + beta·Z over ns partials (ld m, stride
// apart), summed in a fixed order.
static int launch_splitk_reduce(cqr_ctx* c, const double* P, int64_t stride, int ns, int64_t m, int64_t n, double* Z,
                                int64_t ldz, double alpha, double beta) {
  hipEvent_t ev = prof_begin(c);
  const int64_t flat = (m * n + 255) / 256;
  if (flat < 256 && ns >= 16) {
    hipLaunchKernelGGL(splitk_reduce_grouped_kernel, dim3((unsigned)((m * n + 63) / 64)), dim3(256), 0, c->stream,
                       P, stride, ns, m, n, Z, ldz, alpha, beta);
  } else {
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)std::min<int64_t>(2048, flat)), dim3(256), 0, c->stream,
                       P, stride, ns, m, n, Z, ldz, alpha, beta);
  }
  prof_end(c, ev, PK_REDUCE, (double)ns * m * n, 8.0 * (ns + (beta != 0.0 ? 2.0 : 1.0)) * m * n);
  return check_launch(c, "splitk_reduce");
}

template <int ROLE, int PREC = CQR_PREC_F64>
static void launch_role(hipStream_t st, int ta, int tb, dim3 grid, const GemmArgs& g) {
  if (!ta && !tb) hipLaunchKernelGGL((mfma_gemm_f64<false, false, ROLE, PREC>), grid, dim3(256), 0, st, g);
  else if (ta && !tb) hipLaunchKernelGGL((mfma_gemm_f64<true, false, ROLE, PREC>), grid, dim3(256), 0, st, g);
  else if (!ta && tb) hipLaunchKernelGGL((mfma_gemm_f64<false, true, ROLE, PREC>), grid, dim3(256), 0, st, g);
  else hipLaunchKernelGGL((mfma_gemm_f64<true, true, ROLE, PREC>), grid, dim3(256), 0, st, g);
}

static void launch_kind(cqr_ctx* c, int ta, int tb, dim3 grid, const GemmArgs& g, int kind,
                        int prec = CQR_PREC_F64) {
  if (prec == CQR_PREC_F32) {
    // only the two trailing-update products run in reduced precision
    if (kind == PK_GEMM_YTA) launch_role<1, CQR_PREC_F32>(c->stream, ta, tb, grid, g);
    else launch_role<2, CQR_PREC_F32>(c->stream, ta, tb, grid, g);
    return;
  }
  if (prec == CQR_PREC_BF16) {
    if (kind == PK_GEMM_YTA) launch_role<1, CQR_PREC_BF16>(c->stream, ta, tb, grid, g);
    else launch_role<2, CQR_PREC_BF16>(c->stream, ta, tb, grid, g);
    return;
  }
  if (kind == PK_GEMM_YTA) launch_role<1>(c->stream, ta, tb, grid, g);
  else if (kind == PK_GEMM_UPD) launch_role<2>(c->stream, ta, tb, grid, g);
  else if (kind == PK_BGEMM) launch_role<3>(c->stream, ta, tb, grid, g);
  else launch_role<0>(c->stream, ta, tb, grid, g);
}

static int launch_mfma(cqr_ctx* c, int ta, int tb, const GemmArgs& g, int nsplit, int kind,
                       int prec = CQR_PREC_F64) {
  hipEvent_t ev = prof_begin(c);
  dim3 grid((unsigned)((g.m + 127) / 128), (unsigned)((g.n + 127) / 128), (unsigned)nsplit);
  launch_kind(c, ta, tb, grid, g, kind, prec);
  const double out = (double)g.m * g.n;
  const double bytes = 8.0 * ((double)g.m * g.k + (double)g.k * g.n +
                              (nsplit > 1 ? nsplit * out : (g.beta != 0.0 ? 2.0 : 1.0) * out));
  prof_end(c, ev, kind, 2.0 * g.m * g.n * g.k, bytes);
  return check_launch(c, "mfma_gemm_f64");
}

// Split-K choice: enough blocks to fill 256 CUs twice, chunks of ≥ 256 rows.
static int choose_splits(int64_t m, int64_t n, int64_t k, int64_t* kchunk) {
  const int64_t tiles = ((m + 127) / 128) * ((n + 127) / 128);
  // one or two output tiles (the next panel's Yᵀ·A): 128-deep chunks, so
  // that the splits alone give every CU a workgroup
  const int64_t kmin = tiles <= 2 ? 128 : 256;
  int64_t s = 1;
  if (tiles < 512 && k > 512) {
    s = (512 + tiles - 1) / tiles;
    s = std::min<int64_t>(s, (k + kmin - 1) / kmin);
    s = std::max<int64_t>(s, 1);
  }
  int64_t kc = (k + s - 1) / s;
  kc = (kc + 15) / 16 * 16;
  s = (k + kc - 1) / kc;
  *kchunk = kc;
  return (int)s;
}

// Small products (m, n, k ≤ 256: the b×b triangles of T, W, L⁻ᵀ and the
// b-wide Z' = Tᵀ·Z): one wave per 16×16 output tile, operands read straight
// from memory (they sit in L2) with the structural masks applied per
// element on the stored coordinates, FP64 MFMA with four accumulator
// chains.  A 128×128×128 product is 64 single-wave workgroups instead of
// one 256-thread tile walking eight LDS slabs.
struct SmallArgs {
  int m, n, k, ta, tb, smA, smB;
  double alpha, beta;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
};

__global__ void __launch_bounds__(64) small_gemm_kernel(SmallArgs g) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  auto x = [&](int r, int kk) {   // op(A)(i0 + r, kk)
    const int i = i0 + r;
    const bool ok = i < g.m && kk < g.k;
    const int64_t sr = g.ta ? kk : i, sc = g.ta ? i : kk;
    const double v = g.A[ok ? sr + sc * g.lda : 0];
    return ok ? gmask(v, sr, sc, g.smA) : 0.0;
  };
  auto y = [&](int kk, int cc) {  // op(B)(kk, j0 + cc)
    const int j = j0 + cc;
    const bool ok = j < g.n && kk < g.k;
    const int64_t sr = g.tb ? j : kk, sc = g.tb ? kk : j;
    const double v = g.B[ok ? sr + sc * g.ldb : 0];
    return ok ? gmask(v, sr, sc, g.smB) : 0.0;
  };
  const dbl4 acc = mfma_tile16_acc4(x, y, 0, g.k, lane);
  const int i = i0 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int j = j0 + (lane >> 4) + 4 * v;
    if (i < g.m && j < g.n) {
      double* cp = g.C + i + (int64_t)j * g.ldc;
      *cp = g.alpha * acc[v] + (g.beta != 0.0 ? g.beta * *cp : 0.0);
    }
  }
}

static int launch_small(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, int smA, const double* B, int64_t ldb, int smB, double beta,
                        double* C, int64_t ldc, int kind) {
  SmallArgs g{};
  g.m = (int)m; g.n = (int)n; g.k = (int)k; g.ta = ta; g.tb = tb; g.smA = smA; g.smB = smB;
  g.alpha = alpha; g.beta = beta; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(small_gemm_kernel, dim3((unsigned)((m + 15) / 16), (unsigned)((n + 15) / 16)), dim3(64), 0,
                     c->stream, g);
  prof_end(c, ev, kind, 2.0 * m * n * k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n));
  return check_launch(c, "small_gemm");
}

// Products with few columns (n ≤ 16; m, k ≤ 256: the solves' R blocks and
// the b×b×nrhs steps of the Q application): one workgroup, op(B) staged in
// LDS, one thread per output row with NR accumulators, reading its row of
// op(A) straight from memory 64 values at a time (all 64 loads in flight
// before the first is used).  Every element of B is read before any element
// of C is written, so C may alias B.
template <int NR>
__global__ void __launch_bounds__(256) small_thin_kernel(SmallArgs g) {
  constexpr int KC = 64;
  __shared__ double Bs[256][NR + 1];
  const int t = threadIdx.x;
  {  // Bs[kk][j] = op(B)(kk, j): thread t loads row kk = t
    double v[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const bool ok = t < g.k && j < g.n;
      const int64_t sr = g.tb ? j : t, sc = g.tb ? t : j;
      v[j] = g.B[ok ? sr + sc * g.ldb : 0];
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const bool ok = t < g.k && j < g.n;
      const int64_t sr = g.tb ? j : t, sc = g.tb ? t : j;
      if (t < g.k) Bs[t][j] = ok ? gmask(v[j], sr, sc, g.smB) : 0.0;
    }
  }
  __syncthreads();
  if (t >= g.m) return;
  double acc[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) acc[j] = 0.0;
  for (int k0 = 0; k0 < g.k; k0 += KC) {
    double v[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int kk = k0 + q;
      const bool ok = kk < g.k;
      const int64_t sr = g.ta ? kk : t, sc = g.ta ? t : kk;
      v[q] = g.A[ok ? sr + sc * g.lda : 0];
    }
#pragma unroll
    for (int q = 0; q < KC; ++q) {
      const int kk = k0 + q;
      if (kk < g.k) {
        const int64_t sr = g.ta ? kk : t, sc = g.ta ? t : kk;
        const double a = gmask(v[q], sr, sc, g.smA);
#pragma unroll
        for (int j = 0; j < NR; ++j) acc[j] += a * Bs[kk][j];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    if (j >= g.n) continue;
    double* cp = g.C + t + (int64_t)j * g.ldc;
    *cp = g.alpha * acc[j] + (g.beta != 0.0 ? g.beta * *cp : 0.0);
  }
}

int launch_small_thin(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                      int64_t lda, int smA, const double* B, int64_t ldb, int smB, double beta, double* C,
                      int64_t ldc, int kind) {
  if (m <= 0 || n <= 0) return CQR_OK;
  if (m > 256 || k > 256 || n > 16)
    return set_err(c, CQR_ERR_INVALID, "small_thin: %lld×%lld×%lld outside 256×16×256", (long long)m, (long long)n,
                   (long long)k);
  SmallArgs g{};
  g.m = (int)m; g.n = (int)n; g.k = (int)k; g.ta = ta; g.tb = tb; g.smA = smA; g.smB = smB;
  g.alpha = alpha; g.beta = beta; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  hipEvent_t ev = prof_begin(c);
  if (n == 1) hipLaunchKernelGGL(small_thin_kernel<1>, dim3(1), dim3(256), 0, c->stream, g);
  else if (n <= 4) hipLaunchKernelGGL(small_thin_kernel<4>, dim3(1), dim3(256), 0, c->stream, g);
  else hipLaunchKernelGGL(small_thin_kernel<16>, dim3(1), dim3(256), 0, c->stream, g);
  prof_end(c, ev, kind, 2.0 * m * n * k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n));
  return check_launch(c, "small_thin");
}

// Generic C := alpha·op(A)·op(B) + beta·C; split-K through the workspace
// when the output is small and K long.  `scratch` must hold nsplit·m·n
// doubles when splitting (callers size it with gemm_f64_scratch).
static int gemm_f64_impl(cqr_ctx* c, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t lda, int smA, const double* B, int64_t ldb, int smB,
                         double beta, double* C, int64_t ldc, double* scratch, int kind = PK_GEMM_SMALL,
                         int prec = CQR_PREC_F64) {
  if (m <= 0 || n <= 0) return CQR_OK;
  if (prec == CQR_PREC_F64 && m <= 256 && n <= 16 && k <= 256)
    return launch_small_thin(c, ta, tb, m, n, k, alpha, A, lda, smA, B, ldb, smB, beta, C, ldc, kind);
  if (prec == CQR_PREC_F64 && m <= 256 && n <= 256 && k <= 256)
    return launch_small(c, ta, tb, m, n, k, alpha, A, lda, smA, B, ldb, smB, beta, C, ldc, kind);
  GemmArgs g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.smA = smA; g.B = B; g.ldb = ldb; g.smB = smB;
  int64_t kc;
  const int ns = scratch ? choose_splits(m, n, k, &kc) : 1;
  if (ns <= 1) {
    g.C = C; g.ldc = ldc; g.kchunk = std::max<int64_t>(k, 1); g.split_stride = 0;
    return launch_mfma(c, ta, tb, g, 1, kind, prec);
  }
  g.C = scratch; g.ldc = m; g.kchunk = kc; g.split_stride = m * n;
  CQR_TRY(launch_mfma(c, ta, tb, g, ns, kind, prec));
  return launch_splitk_reduce(c, scratch, m * n, ns, m, n, C, ldc, alpha, beta);
}

// Batched over the problems of a group: operand modes BM_PROB_* and
// BM_STRIDED; m_prob selects per-problem rows (m of problem z − prow).
int gemm_f64_batched_prob(cqr_ctx* c, int ta, int tb, int m, int n, int k, double alpha, const BMat& A, int smA,
                          const BMat& B, int smB, double beta, const BMat& C, int smC, int nprob, const BatchProb* bp,
                          int64_t prow, int m_prob, int max_m) {
  if (nprob <= 0 || n <= 0 || max_m <= 0) return CQR_OK;
  GemmArgs g{};
  g.m = m_prob ? max_m : m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A.p; g.lda = A.ld; g.smA = smA; g.B = B.p; g.ldb = B.ld; g.smB = smB; g.C = const_cast<double*>(C.p);
  g.ldc = C.ld; g.kchunk = k; g.split_stride = 0;
  g.batched = 1;
  g.a_mode = A.mode; g.b_mode = B.mode; g.c_mode = C.mode;
  g.a_zs = A.zs; g.b_zs = B.zs; g.c_zs = C.zs;
  g.smC = smC; g.bp = bp; g.prow = prow; g.m_prob = m_prob;
  dim3 grid((unsigned)((max_m + 127) / 128), (unsigned)((n + 127) / 128), (unsigned)nprob);
  hipEvent_t ev = prof_begin(c);
  launch_kind(c, ta, tb, grid, g, PK_GEMM_SMALL);
  prof_end(c, ev, PK_GEMM_SMALL, 0.0, 0.0);
  return check_launch(c, "mfma_gemm_f64_batched_prob");
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

int lean_gemm(cqr_ctx* c, int ak, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* scratch, int nsplit,
              int kind, int* used);


// FP64 update with the unmasked bulk on the lean kernel: rows [0, bt) of Y
// (its unit-lower top block) go through the masked general kernel, rows
// [bt, m) through lean_gemm; the Yᵀ·A partial sums of both are reduced in a
// fixed order.
static int wy_update_f64_split(cqr_ctx* c, const double* Y, int64_t ldy, int smY, const double* Tt, int t_trans,
                               double* A, int64_t lda, int64_t m, int64_t n, int b, double* part, int ns, double* Z,
                               double* Zp) {
  const int64_t bt = std::min<int64_t>(b, m);
  // slot 0: Y(0:bt)ᵀ·A(0:bt), masked
  GemmArgs g{};
  g.m = b; g.n = n; g.k = bt; g.alpha = 1.0; g.beta = 0.0;
  g.A = Y; g.lda = ldy; g.smA = smY; g.B = A; g.ldb = lda; g.smB = SM_NONE;
  g.C = part; g.ldc = b; g.kchunk = std::max<int64_t>(bt, 1); g.split_stride = (int64_t)b * n;
  CQR_TRY(launch_mfma(c, 1, 0, g, 1, PK_GEMM_YTA));
  // slots 1..: Y(bt:m)ᵀ·A(bt:m) on the lean kernel
  int used = 0;
  CQR_TRY(lean_gemm(c, 1, b, n, m - bt, 1.0, Y + bt, ldy, A + bt, lda, 0.0, nullptr, 0, part + (int64_t)b * n,
                    ns - 1, PK_GEMM_YTA, &used));
  const int nslots = 1 + used;
  CQR_TRY(launch_splitk_reduce(c, part, (int64_t)b * n, nslots, b, n, Z, b, 1.0, 0.0));
  CQR_TRY(gemm_f64_impl(c, t_trans ? 1 : 0, 0, b, n, b, 1.0, Tt, b, SM_LOWER, Z, b, SM_NONE, 0.0, Zp, b, nullptr));
  // A(0:bt) −= Y(0:bt)·Z' (masked), A(bt:m) −= Y(bt:m)·Z' (lean)
  CQR_TRY(gemm_f64_impl(c, 0, 0, bt, n, b, -1.0, Y, ldy, smY, Zp, b, SM_NONE, 1.0, A, lda, nullptr, PK_GEMM_UPD));
  return update_gemm(c, m - bt, n, b, -1.0, Y + bt, ldy, Zp, b, 1.0, A + bt, lda, PK_GEMM_UPD, 0);
}

int wy_update_ws(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt, int t_trans,
                 double* A, int64_t lda, int64_t m, int64_t n, int b, double* ws, int prec) {
  if (m <= 0 || n <= 0 || b <= 0) return CQR_OK;
  const int smY = y_implicit ? SM_UNIT_LOWER : SM_NONE;
  double* part = ws;
  const size_t np = gemm_f64_scratch_doubles(b, n, m);
  double* Z = ws + ((np + 31) / 32) * 32;
  double* Zp = Z + ((size_t)b * n + 31) / 32 * 32;
  // the lean kernel's 128×128 tiles pay off from three column tiles on
  if (prec == CQR_PREC_F64 && np > 0 && m > 2 * b && n > 2 * 128) {
    const int ns = (int)(np / ((size_t)b * n));   // partial slots of choose_splits
    if (ns >= 2) return wy_update_f64_split(c, Y, ldy, smY, Tt, t_trans, A, lda, m, n, b, part, ns, Z, Zp);
  }
  // Z = Yᵀ·A  (b×n)
  CQR_TRY(gemm_f64_impl(c, 1, 0, b, n, m, 1.0, Y, ldy, smY, A, lda, SM_NONE, 0.0, Z, b,
                        np ? part : nullptr, PK_GEMM_YTA, prec));
  const bool plain = prec == CQR_PREC_F64 && m > 2 * b;
  // Z' = Tᵀ·Z (Qᵀ·A), or T·Z with t_trans (Q·A); Tt holds Tᵀ (lower), FP64
  CQR_TRY(gemm_f64_impl(c, t_trans ? 1 : 0, 0, b, n, b, 1.0, Tt, b, SM_LOWER, Z, b, SM_NONE, 0.0, Zp, b,
                        nullptr));
  // A −= Y·Z' (FP64: the top b rows, where Y may be unit lower, on the
  // masked kernel, the rest on the rank-b update kernel)
  if (plain) {
    CQR_TRY(gemm_f64_impl(c, 0, 0, b, n, b, -1.0, Y, ldy, smY, Zp, b, SM_NONE, 1.0, A, lda, nullptr, PK_GEMM_UPD));
    return update_gemm(c, m - b, n, b, -1.0, Y + b, ldy, Zp, b, 1.0, A + b, lda, PK_GEMM_UPD, 0);
  }
  return gemm_f64_impl(c, 0, 0, m, n, b, -1.0, Y, ldy, smY, Zp, b, SM_NONE, 1.0, A, lda, nullptr,
                       PK_GEMM_UPD, prec);
}

// Trailing update with an explicit unit-lower Y (Ye, m×b, ld m) and
// YT = Ye·Tᵀ-factor product (m×b, ld m, i.e. Ye·Tt): A (m×n) −= YT·(Yeᵀ·A).
// No structural masks remain, so both products are plain: Z = Yeᵀ·A on the
// lean kernel (split-K, from three column tiles on) or the general one, and
// the update through the vendor BLAS.
// Z (b×n, in ws) = Yᵀ·A: the first product of the explicit update.
int wy_explicit_yta(cqr_ctx* c, const double* Ye, int64_t m, int64_t n, int b, const double* A, int64_t lda,
                    double* ws, int prec, double** Zout) {
  double* part = ws;
  const size_t np = gemm_f64_scratch_doubles(b, n, m);
  double* Z = ws + ((np + 31) / 32) * 32;
  *Zout = Z;
  if (m <= 0 || n <= 0 || b <= 0) return CQR_OK;
  if (prec != CQR_PREC_F64)   // FP32/BF16 operands, FP32 accumulation; Z lands in FP64
    return gemm_f64_impl(c, 1, 0, b, n, m, 1.0, Ye, m, SM_NONE, A, lda, SM_NONE, 0.0, Z, b, np ? part : nullptr,
                         PK_GEMM_YTA, prec);
  const int ns = np ? (int)(np / ((size_t)b * n)) : 0;
  if (ns >= 2) {
    int used = 0;
    CQR_TRY(lean_gemm(c, 1, b, n, m, 1.0, Ye, m, A, lda, 0.0, nullptr, 0, part, ns, PK_GEMM_YTA, &used));
    return launch_splitk_reduce(c, part, (int64_t)b * n, used, b, n, Z, b, 1.0, 0.0);
  }
  return gemm_f64_impl(c, 1, 0, b, n, m, 1.0, Ye, m, SM_NONE, A, lda, SM_NONE, 0.0, Z, b, np ? part : nullptr,
                       PK_GEMM_YTA);
}

// A −= (Y·Tᵀ)·Z: the second product of the explicit update.
int wy_explicit_apply(cqr_ctx* c, const double* YT, int64_t m, int64_t n, int b, const double* Z, double* A,
                      int64_t lda, int prec) {
  if (m <= 0 || n <= 0 || b <= 0) return CQR_OK;
  if (prec != CQR_PREC_F64)
    return gemm_f64_impl(c, 0, 0, m, n, b, -1.0, YT, m, SM_NONE, Z, b, SM_NONE, 1.0, A, lda, nullptr, PK_GEMM_UPD,
                         prec);
  return update_gemm(c, m, n, b, -1.0, YT, m, Z, b, 1.0, A, lda, PK_GEMM_UPD, 0);
}

int wy_update_explicit(cqr_ctx* c, const double* Ye, const double* YT, int64_t m, int64_t n, int b, double* A,
                       int64_t lda, double* ws, int prec) {
  if (m <= 0 || n <= 0 || b <= 0) return CQR_OK;
  double* Z = nullptr;
  CQR_TRY(wy_explicit_yta(c, Ye, m, n, b, A, lda, ws, prec, &Z));
  return wy_explicit_apply(c, YT, m, n, b, Z, A, lda, prec);
}

int wy_update_f64_ws(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt,
                     double* A, int64_t lda, int64_t m, int64_t n, int b, double* ws) {
  return wy_update_ws(c, Y, ldy, y_implicit, Tt, 0, A, lda, m, n, b, ws, CQR_PREC_F64);
}

int wy_update_f64(cqr_ctx* c, const double* Y, int64_t ldy, int y_implicit, const double* Tt,
                  double* A, int64_t lda, int64_t m, int64_t n, int b) {
  const size_t need = wy_update_ws_doubles(m, n, b);
  CQR_TRY(ensure_ws(c, need * sizeof(double)));
  return wy_update_f64_ws(c, Y, ldy, y_implicit, Tt, A, lda, m, n, b, static_cast<double*>(c->ws));
}

}  // namespace cqr
