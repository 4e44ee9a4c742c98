// Panel kernels of the CA-QR hot path for gfx950 (FP64):
//   leaf_qr       K1  Householder QR of every leaf row block (dgeqrf semantics)
//   tree_node     K2  QR of stacked triangles [R_top; R_bot] (dtpqrt l=b)
//   bgemm         small batched GEMMs for the reverse tree, the per-leaf
//                 block-reflector products and the Y formation
//   hr_core       K4  LU without pivoting of Q1_top − S, W, U⁻¹, L⁻ᵀ
// plus element-wise helpers.  Leaves are contiguous row blocks
// [floor(l·m/L), floor((l+1)·m/L)); the tree pairs receiver r with sender
// r + ceil(n/2) on a level of n live ranks (alg/QR/tsqr/bitree_tsqr.cxx:172-208),
// so every address is computed on the device from scalar arguments.
#include "cqr_device.h"
#include "cqr_internal.h"
#include "leaf_body.h"

namespace cqr {


// Loop over the entries (rr, cc) of a b×b matrix: waves over columns, lanes
// over rows (no integer division; coalesced along columns).
#define FOR_RC(rr, cc, b, NT)                                                     \
  for (int cc = wave_id(); cc < (b); cc += (NT) / 64)                             \
    for (int rr = (int)(threadIdx.x & 63); rr < (b); rr += 64)

// ---------------------------------------------------------------------------
// K1: leaf QR, one workgroup per leaf (body in leaf_body.h).
// ---------------------------------------------------------------------------
template <int WAVES, int CPW, int RPL>
__global__ void __launch_bounds__(64 * WAVES)
leaf_qr_kernel(double* __restrict__ A, int64_t lda, int64_t m, int L, int b,
               double* __restrict__ Rout, double* __restrict__ Tout, double* __restrict__ tauout,
               const BatchProb* __restrict__ bp, int nprob) {
  extern __shared__ double smem[];
  leaf_qr_body<WAVES, CPW, RPL>(A, lda, m, L, b, Rout, Tout, tauout, bp, nprob, blockIdx.x, smem, nullptr, 0);
}

// Packed column-major triangles in LDS: upper (r ≤ c) and strictly upper.
__device__ __forceinline__ int pu(int r, int c) { return (c * (c + 1) >> 1) + r; }
__device__ __forceinline__ int psu(int r, int c) { return (c * (c - 1) >> 1) + r; }

// ---------------------------------------------------------------------------
// K2: one tree level.  Node z stacks R of receiver z on R of sender
// z + ceil(n/2) and computes its QR like dtpqrt with l = b: the reflector
// for column i is [e_i; x] with x = R_bot(0:i, i), so R_bot becomes V2
// (upper triangular, in registers as in K1) and R_top (packed in LDS)
// becomes the new R of the receiver.  Row i of R_top couples into column
// j's dot product only through its entry (i, j): the lane that ends up
// holding column j's reduced sum adds it and applies the update.
// ---------------------------------------------------------------------------
template <int WAVES, int CPW, int RPL>
__global__ void __launch_bounds__(64 * WAVES)
tree_node_kernel(double* __restrict__ Rleaf, int n, int node0, int b, double* __restrict__ V2out,
                 double* __restrict__ Tout, double* __restrict__ tauout) {
  constexpr int NT = 64 * WAVES;
  const int z = blockIdx.x;
  const int snd = z + (n + 1) / 2;
  const int64_t bb = (int64_t)b * b;
  double* Rt = Rleaf + z * bb;
  const double* Rb = Rleaf + snd * bb;
  double* V2 = V2out + (node0 + z) * bb;
  double* Tn = Tout + (node0 + z) * bb;
  double* taun = tauout + (int64_t)(node0 + z) * b;

  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  extern __shared__ double smem[];
  const int nu = b * (b + 1) / 2, nsu = b * (b - 1) / 2;
  double* R1 = smem;                      // packed upper b×b
  double* Gp = R1 + nu;                   // packed strictly-upper Gram
  double* vbuf = Gp + nsu;                // 2 × 64·RPL
  double* stau = vbuf + 2 * 64 * RPL;     // b

  double a2[CPW][RPL];
#pragma unroll
  for (int cl = 0; cl < CPW; ++cl) {
    const int col = w + WAVES * cl;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      a2[cl][s] = (col < b && row <= col) ? Rb[row + col * b] : 0.0;
    }
  }
  for (int cc = w; cc < b; cc += WAVES) {
    for (int rr = lane; rr <= cc; rr += 64) R1[pu(rr, cc)] = Rt[rr + cc * b];
    for (int rr = lane; rr < cc; rr += 64) Gp[psu(rr, cc)] = 0.0;
  }
  __syncthreads();

  for (int i = 0; i < b; ++i) {
    const int ow = i % WAVES, ocl = i / WAVES;
    double* vb = vbuf + (i & 1) * (64 * RPL);
    if (w == ow) {
#pragma unroll
      for (int cl = 0; cl < CPW; ++cl) {
        if (cl != ocl) continue;
        double ss = 0.0;
#pragma unroll
        for (int s = 0; s < RPL; ++s)
          if (lane + 64 * s <= i) ss += a2[cl][s] * a2[cl][s];
        ss = wave_sum(ss);
        const double alpha = R1[pu(i, i)];
        double xnorm;
        if (ss_needs_rescale(ss)) {
          double mx = 0.0;
#pragma unroll
          for (int s = 0; s < RPL; ++s)
            if (lane + 64 * s <= i) mx = fmax(mx, fabs(a2[cl][s]));
          mx = wave_max(mx);
          if (mx == 0.0) {
            xnorm = 0.0;
          } else {
            double s2 = 0.0;
#pragma unroll
            for (int s = 0; s < RPL; ++s)
              if (lane + 64 * s <= i) { const double t = a2[cl][s] / mx; s2 += t * t; }
            xnorm = mx * sqrt(wave_sum(s2));
          }
        } else {
          xnorm = sqrt(ss);
        }
        const Reflector h = dlarfg_scalar(alpha, xnorm);
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
          const int row = lane + 64 * s;
          if (h.tau != 0.0 && row <= i) {
            double x = a2[cl][s];
            for (int k = 0; k < h.knt; ++k) x *= 0x1p969;
            a2[cl][s] = x * h.scal;
          }
          vb[row] = row <= i ? a2[cl][s] : 0.0;
        }
        if (lane == 0) {
          R1[pu(i, i)] = h.beta;
          stau[i] = h.tau;
        }
      }
    }
    __syncthreads();
    const double tau = stau[i];
    double v[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) v[s] = vb[lane + 64 * s];
    double part[CPW];
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s) acc += v[s] * a2[cl][s];
      part[cl] = acc;
    }
    double d = xreduce<CPW>(part, lane);
    {
      const int col = w + WAVES * (lane & (CPW - 1));
      if (lane < CPW && col < b) {
        if (col > i) {
          const double r1 = R1[pu(i, col)];
          d += r1;
          R1[pu(i, col)] = r1 - tau * d;
        } else if (col < i) {
          Gp[psu(col, i)] = d;
        }
      }
    }
#pragma unroll
    for (int cl = 0; cl < CPW; ++cl) {
      const int col = w + WAVES * cl;
      if (col > i && col < b) {
        const double f = tau * readlane_d(d, cl);
#pragma unroll
        for (int s = 0; s < RPL; ++s) a2[cl][s] -= f * v[s];
      }
    }
  }
  __syncthreads();
  // outputs: new R of the receiver, V2; then T from the Gram in dense LDS
  for (int cc = w; cc < b; cc += WAVES)
    for (int rr = lane; rr < b; rr += 64) Rt[rr + cc * b] = rr <= cc ? R1[pu(rr, cc)] : 0.0;
#pragma unroll
  for (int cl = 0; cl < CPW; ++cl) {
    const int col = w + WAVES * cl;
    if (col >= b) continue;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int row = lane + 64 * s;
      if (row < b) V2[row + col * b] = row <= col ? a2[cl][s] : 0.0;
    }
  }
  // Gram (packed, after R1) → dense b×b at the start of LDS, via registers
  constexpr int GPT = (CQR_MAX_B * (CQR_MAX_B - 1) / 2 + NT - 1) / NT;
  double g[GPT];
#pragma unroll
  for (int t = 0; t < GPT; ++t) {
    const int e = tid + t * NT;
    g[t] = e < nsu ? Gp[e] : 0.0;
  }
  __syncthreads();
  double* M = smem;
  for (int cc = w; cc < b; cc += WAVES)
    for (int rr = lane; rr < b; rr += 64) M[rr + cc * b] = 0.0;
  __syncthreads();
  // packed strictly-upper index e → (r, c): column c holds c entries
#pragma unroll
  for (int t = 0; t < GPT; ++t) {
    const int e = tid + t * NT;
    if (e < nsu) {
      int c = (int)((1.0 + sqrt(1.0 + 8.0 * e)) * 0.5);
      while (c * (c - 1) / 2 > e) --c;
      while ((c + 1) * c / 2 <= e) ++c;
      const int rr = e - c * (c - 1) / 2;
      M[rr + c * b] = g[t];
    }
  }
  __syncthreads();
  tri_inv_upper_lds<NT>(M, b, stau, b);
  for (int cc = w; cc < b; cc += WAVES)
    for (int rr = lane; rr < b; rr += 64) Tn[rr + cc * b] = rr <= cc ? M[rr + cc * b] : 0.0;
  for (int p = tid; p < b; p += NT) taun[p] = stau[p];
}

// ---------------------------------------------------------------------------
// Batched small GEMM (FP64 FMA, 64×64 tiles, 256 threads, 4×4 per thread):
//   C_z := alpha·op(A_z)·op(B_z) + beta·C_z
// Operands are addressed per batch index z (see BMat), carry structural
// masks in stored coordinates, and the number of rows may follow the leaf
// partition of a panel (m < 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ const double* bmat_ptr(const BMat& M, int z, const BgemmArgs& g) {
  switch (M.mode) {
    case BM_STRIDED: return M.p + (int64_t)z * M.zs;
    case BM_LEAFROWS: return M.p + leaf_off(z, g.pm, g.pL);
    case BM_LEAFROWS_SHIFTED: return M.p + leaf_off(z + g.zoff, g.pm, g.pL);
    default: return M.p + (int64_t)(z + g.zoff) * M.zs;
  }
}

__device__ __forceinline__ double apply_mask(double v, int64_t r, int64_t c, int sm) {
  switch (sm) {
    case SM_UNIT_LOWER: return r < c ? 0.0 : (r == c ? 1.0 : v);
    case SM_UPPER: return r > c ? 0.0 : v;
    case SM_LOWER: return r < c ? 0.0 : v;
    case SM_IDENTITY: return r == c ? 1.0 : 0.0;
    case SM_UNIT_UPPER: return r > c ? 0.0 : (r == c ? 1.0 : v);
    default: return v;
  }
}

__global__ void __launch_bounds__(256) bgemm_kernel(BgemmArgs g) {
  const int z = blockIdx.z;
  int m = g.m;
  if (m < 0) m = (int)(leaf_off(z + 1, g.pm, g.pL) - leaf_off(z, g.pm, g.pL));
  int k = g.k;
  if (k < 0) k = (int)(leaf_off(z + 1, g.pm, g.pL) - leaf_off(z, g.pm, g.pL));
  const int n = g.n;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  if (m0 >= m || n0 >= n) return;
  const double* A = bmat_ptr(g.A, z, g);
  const double* B = bmat_ptr(g.B, z, g);
  double* C = const_cast<double*>(bmat_ptr(g.C, z, g));
  const int64_t lda = g.A.ld, ldb = g.B.ld, ldc = g.C.ld;

  // 64×64 output tile, 16-deep k slabs in LDS ([k][m] and [k][n]); each of
  // the four waves owns a 32×32 quarter as 2×2 FP64 MFMA 16×16 tiles
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int fr = lane >> 4, fc = lane & 15;
  const int wm0 = (wave & 1) * 32, wn0 = (wave >> 1) * 32;
  dbl4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < k; k0 += 16) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int e = tid + 256 * rr;
      int mm, kk;
      if (!g.ta) { mm = e & 63; kk = e >> 6; } else { kk = e & 15; mm = e >> 4; }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < m && gk < k) {
        const int64_t sr = g.ta ? gk : gm, sc = g.ta ? gm : gk;
        v = (g.smA == SM_IDENTITY) ? 0.0 : A[sr + sc * lda];
        v = apply_mask(v, sr, sc, g.smA);
      }
      As[kk][mm] = v;
      int nn;
      if (!g.tb) { kk = e & 15; nn = e >> 4; } else { nn = e & 63; kk = e >> 6; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      double u = 0.0;
      if (gn < n && gk2 < k) {
        const int64_t sr = g.tb ? gn : gk2, sc = g.tb ? gk2 : gn;
        u = (g.smB == SM_IDENTITY) ? 0.0 : B[sr + sc * ldb];
        u = apply_mask(u, sr, sc, g.smB);
      }
      Bs[kk][nn] = u;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = As[kk + fr][wm0 + 16 * i + fc];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Bs[kk + fr][wn0 + 16 * j + fc];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // lane l, register v of acc[i][j] holds C[m0+wm0+16i+(l&15)][n0+wn0+16j+(l>>4)+4v]
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int gm = m0 + wm0 + 16 * i + fc;
    if (gm >= m) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gn = n0 + wn0 + 16 * j + fr + 4 * v;
        if (gn >= n) continue;
        double* cp = C + gm + (int64_t)gn * ldc;
        double val = g.alpha * acc[i][j][v];
        if (g.beta != 0.0) val += g.beta * *cp;
        if (g.smC == SM_UPPER && gm > gn) val = 0.0;
        *cp = val;
      }
  }
}

// ---------------------------------------------------------------------------
// Signed LU of a b×b block (alg/QR/hh_recon/hh_recon.cxx:43-70), b ≤ 128,
// 512 threads, in LDS (M, ld b).  Step i picks s_i (−1 if the pivot is
// positive, else +1), subtracts it from the pivot, scales the column below
// the pivot by its reciprocal (|pivot| ≥ 1) and updates the later columns.
// Blocked by eight: one wave factors the block's columns from registers
// (pivot-row entries by readlane, no barrier); one thread per later column
// solves its eight block rows with the unit-lower diagonal block (U12);
// the rank-8 update of the trailing matrix runs as FP64 MFMA tiles over all
// waves.  On exit M holds L\U and sgn the signs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void signed_lu_lds(const double* __restrict__ src, int64_t ld, double* M, int ldm, double* sgn, int b) {
  constexpr int NT = 512, NB = 8, RPL = 2;
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  {
    // all 32 loads of a thread in flight at once (clamped addresses, then masked stores)
    double v[16][2];
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        v[j][s] = src[(cc < b && rr < b) ? rr + (int64_t)cc * ld : 0];
      }
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        if (cc < b && rr < b) M[rr + cc * ldm] = v[j][s];
      }
  }
  __syncthreads();
  for (int c0 = 0; c0 < b; c0 += NB) {
    const int nb = min(NB, b - c0);
    if (w == 0) {
      double pn[RPL][NB];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int row = lane + 64 * s;
#pragma unroll
        for (int j = 0; j < NB; ++j) pn[s][j] = (row >= c0 && row < b && j < nb) ? M[row + (c0 + j) * ldm] : 0.0;
      }
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        if (jj < nb) {
        const int i = c0 + jj, hi = i >> 6, pl = i & 63;
        const double pv = readlane_d(hi ? pn[1][jj] : pn[0][jj], pl);
        const double sg = pv > 0.0 ? -1.0 : 1.0;
        const double piv = pv - sg;
        const double rp = 1.0 / piv;
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
          const int row = lane + 64 * s;
          if (row > i) pn[s][jj] *= rp;
          if (row == i) pn[s][jj] = piv;
        }
        if (lane == 0) sgn[i] = sg;
#pragma unroll
        for (int j2 = jj + 1; j2 < NB; ++j2) {
          if (j2 < nb) {
            const double u = readlane_d(hi ? pn[1][j2] : pn[0][j2], pl);
#pragma unroll
            for (int s = 0; s < RPL; ++s)
              if (lane + 64 * s > i) pn[s][j2] -= pn[s][jj] * u;
          }
        }
        }
      }
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int row = lane + 64 * s;
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (row >= c0 && row < b && j < nb) M[row + (c0 + j) * ldm] = pn[s][j];
      }
    }
    __syncthreads();
    const int j0 = c0 + nb;
    if (j0 >= b) break;
    // U12 = L11⁻¹·A12, one thread per later column
    for (int j = j0 + tid; j < b; j += NT) {
      double u[NB];
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) u[jj] = jj < nb ? M[c0 + jj + j * ldm] : 0.0;
#pragma unroll
      for (int jj = 1; jj < NB; ++jj)
#pragma unroll
        for (int p = 0; p < jj; ++p)
          if (jj < nb) u[jj] -= M[c0 + jj + (c0 + p) * ldm] * u[p];
#pragma unroll
      for (int jj = 1; jj < NB; ++jj)
        if (jj < nb) M[c0 + jj + j * ldm] = u[jj];
    }
    __syncthreads();
    // A22 −= L21·U12 on 16×16 MFMA tiles (K = 8)
    const int nt = (b - j0 + 15) / 16;
    auto tile = [&](int t) {
      const int r0 = j0 + 16 * (t % nt), q0 = j0 + 16 * (t / nt);
      auto x = [&](int m, int k) {
        const int r = r0 + m;
        const bool ok = r < b && k < nb;
        return ok ? M[r + (c0 + k) * ldm] : 0.0;
      };
      auto y = [&](int k, int n) {
        const int q = q0 + n;
        const bool ok = q < b && k < nb;
        return ok ? M[c0 + k + q * ldm] : 0.0;
      };
      const dbl4 acc = mfma_tile16(x, y, 0, NB, lane);
      const int r = r0 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = q0 + (lane >> 4) + 4 * v;
        if (r < b && q < b) M[r + q * ldm] -= acc[v];
      }
    };
    // two tiles per wave in flight
    for (int t = w; t < nt * nt; t += 2 * NT / 64) {
      tile(t);
      if (t + NT / 64 < nt * nt) tile(t + NT / 64);
    }
    __syncthreads();
  }
  __syncthreads();
}

// K4: Householder reconstruction core (one workgroup).  M is kept in LDS
// with leading dimension b + 1.
__device__ __forceinline__ void hr_core_body(const double* __restrict__ Q1top, int64_t ldq, const double* __restrict__ Rroot,
                             double* __restrict__ Atop, int64_t lda, double* __restrict__ W,
                             double* __restrict__ Uinv, double* __restrict__ Winv, double* __restrict__ LinvT,
                             int64_t* __restrict__ signs, int b) {
  constexpr int NT = 512;
  extern __shared__ double smem[];
  const int ldm = b + 1;
  double* M = smem;            // b×b, ld b+1
  double* sgn = M + b * ldm;   // b
  double* d = sgn + b;         // b
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  signed_lu_lds(Q1top, ldq, M, ldm, sgn, b);
  // outputs: L below the diagonal of Atop, S·R on and above it, W = U·S
  {
    double rv[16][2];
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        rv[j][s] = Rroot[(cc < b && rr < b) ? rr + cc * b : 0];
      }
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        if (cc < b && rr < b) {
          const double mv = M[rr + cc * ldm];
          Atop[rr + cc * lda] = rr > cc ? mv : sgn[rr] * rv[j][s];
          W[rr + cc * b] = rr <= cc ? mv * sgn[cc] : 0.0;
        }
      }
  }
  if (signs)
    for (int p = tid; p < b; p += NT) signs[p] = (int64_t)sgn[p];
  for (int p = tid; p < b; p += NT) d[p] = 1.0 / M[p + p * ldm];
  __syncthreads();
  tri_inv_upper_lds<NT>(M, ldm, d, b);
  FOR_RC(rr, cc, b, NT) {
    const double u = rr <= cc ? M[rr + cc * ldm] : 0.0;
    Uinv[rr + cc * b] = u;
    Winv[rr + cc * b] = sgn[rr] * u;        // W⁻¹ = S·U⁻¹
  }
  __syncthreads();
  // L⁻ᵀ: load Lᵀ (unit upper) from Atop and invert
  {
    double lv[16][2];
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        lv[j][s] = Atop[(cc < b && rr < cc) ? cc + rr * lda : 0];
      }
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int cc = w + 8 * j, rr = lane + 64 * s;
        if (cc < b && rr < b) M[rr + cc * ldm] = rr < cc ? lv[j][s] : 0.0;
      }
  }
  for (int p = tid; p < b; p += NT) d[p] = 1.0;
  __syncthreads();
  tri_inv_upper_lds<NT>(M, ldm, d, b);
  FOR_RC(rr, cc, b, NT) LinvT[rr + cc * b] = rr <= cc ? M[rr + cc * ldm] : 0.0;
}

// The products that follow the HR core in a QR_2D panel, in the same
// workgroup (b = 128, 8 waves; wave w owns the 16-row block w of each
// result).  On entry L⁻ᵀ is in M (upper part, ld b+1) and W, U⁻¹, L are in
// global memory, stored by this workgroup.
//   Tt = Tᵀ = −L⁻¹·Wᵀ  (lower; T = −W·L⁻ᵀ, hh_recon.cxx:282-308)
//   M2 = U⁻¹·Tᵀ        (so that Q1·[U⁻¹ | M2] gives Y and Y·Tᵀ at once)
//   with Ye: Ye top = [unit lower L],  YT top = Y1·Tᵀ = −Wᵀ  (W = −T·Y1ᵀ)
__device__ __forceinline__ void hr_tail_body(double* __restrict__ M, int ldm, const double* __restrict__ Atop,
                                             int64_t lda, const double* __restrict__ W,
                                             const double* __restrict__ Uinv, double* __restrict__ Tt,
                                             double* __restrict__ M2, double* __restrict__ Ye,
                                             double* __restrict__ YT, int64_t ldy) {
  constexpr int b = 128;
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const int r16 = lane & 15, q4 = lane >> 4;
  const int i0 = 16 * w;
  __syncthreads();   // W, U⁻¹ and L stored by every wave are visible
  dbl4 tt[8];
#pragma unroll
  for (int tn = 0; tn < 8; ++tn) {
    tt[tn] = dbl4{0.0, 0.0, 0.0, 0.0};
    if (tn > w) continue;   // Tᵀ is lower: tiles right of the diagonal are zero
    // L⁻¹[i][k] = L⁻ᵀ[k][i], nonzero for k ≤ i;  Wᵀ[k][j] = W[j][k], nonzero for j ≤ k
    auto x = [&](int m, int k) {
      const int i = i0 + m;
      const bool ok = k <= i;
      const double v = M[ok ? k + i * ldm : 0];   // clamped, loaded unconditionally
      return ok ? v : 0.0;
    };
    auto y = [&](int k, int n) {
      const int j = 16 * tn + n;
      const bool ok = j <= k && k < b;
      const double v = W[ok ? j + k * b : 0];
      return ok ? v : 0.0;
    };
    tt[tn] = mfma_tile16_acc4(x, y, 16 * tn, i0 + 16, lane);
  }
  __syncthreads();   // every wave is done reading L⁻ᵀ from M
  // Tᵀ (negated) into M, column-major ld ldm, and out
#pragma unroll
  for (int tn = 0; tn < 8; ++tn)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + r16, j = 16 * tn + q4 + 4 * v;
      const double val = -tt[tn][v];
      M[i + j * ldm] = val;
      Tt[i + j * b] = val;
    }
  __syncthreads();
  if (M2) {
#pragma unroll
    for (int tn = 0; tn < 8; ++tn) {
      // U⁻¹[i][k] nonzero for k ≥ i; Tᵀ[k][j] nonzero for k ≥ j
      auto x = [&](int m, int k) {
        const int i = i0 + m;
        const bool ok = k >= i && k < b;
        const double v = Uinv[ok ? i + k * b : 0];
        return ok ? v : 0.0;
      };
      auto y = [&](int k, int n) {
        const int j = 16 * tn + n;
        const bool ok = k >= j && k < b;
        const double v = M[ok ? k + j * ldm : 0];
        return ok ? v : 0.0;
      };
      const dbl4 acc = mfma_tile16_acc4(x, y, max(i0, 16 * tn), b, lane);
#pragma unroll
      for (int v = 0; v < 4; ++v) M2[(i0 + r16) + (16 * tn + q4 + 4 * v) * b] = acc[v];
    }
  }
  if (Ye) {
    FOR_RC(rr, cc, b, 512) {
      Ye[rr + cc * ldy] = rr > cc ? Atop[rr + cc * lda] : (rr == cc ? 1.0 : 0.0);
      YT[rr + cc * ldy] = rr >= cc ? -W[cc + rr * b] : 0.0;
    }
  }
}

__global__ void __launch_bounds__(512)
hr_core_kernel(const double* __restrict__ Q1top, int64_t ldq, const double* __restrict__ Rroot,
               double* __restrict__ Atop, int64_t lda, double* __restrict__ W,
               double* __restrict__ Uinv, double* __restrict__ Winv, double* __restrict__ LinvT,
               int64_t* __restrict__ signs, int b, double* __restrict__ Tt, double* __restrict__ M2,
               double* __restrict__ Ye, double* __restrict__ YT, int64_t ldy) {
  hr_core_body(Q1top, ldq, Rroot, Atop, lda, W, Uinv, Winv, LinvT, signs, b);
  if (Tt) {
    extern __shared__ double smem[];
    hr_tail_body(smem, b + 1, Atop, lda, W, Uinv, Tt, M2, Ye, YT, ldy);
  }
}

// Batched: workgroup p handles problem p (its Q1 top in Y, its root R).
__global__ void __launch_bounds__(512)
hr_core_batched_kernel(const BatchProb* __restrict__ bp, const double* __restrict__ Y,
                       const double* __restrict__ leafR, const double* __restrict__ nodeR,
                       double* __restrict__ Abase, double* __restrict__ W, double* __restrict__ Uinv,
                       double* __restrict__ Winv, double* __restrict__ LinvT, int b) {
  const BatchProb P = bp[blockIdx.x];
  const int64_t bb = (int64_t)b * b, p = blockIdx.x;
  const double* Rroot = P.L > 1 ? nodeR + (int64_t)(P.node0 + P.L - 2) * bb : leafR + (int64_t)P.leaf0 * bb;
  hr_core_body(Y + P.yoff, P.m, Rroot, Abase + P.aoff, P.m, W + p * bb, Uinv + p * bb, Winv + p * bb,
               LinvT + p * bb, nullptr, b);
}

// a5 standalone: signed LU of a b×b block with R scaling.
__global__ void __launch_bounds__(512)
signed_nlu_kernel(double* __restrict__ A, int64_t lda, double* __restrict__ R, int64_t ldr,
                  int64_t* __restrict__ signs, int b) {
  constexpr int NT = 512;
  extern __shared__ double smem[];
  const int ldm = b + 1;
  double* M = smem;
  double* sgn = M + b * ldm;
  const int tid = threadIdx.x;
  signed_lu_lds(A, lda, M, ldm, sgn, b);
  FOR_RC(rr, cc, b, NT) {
    A[rr + cc * lda] = M[rr + cc * ldm];
    if (R && rr <= cc && sgn[rr] < 0) R[rr + cc * ldr] = -R[rr + cc * ldr];
  }
  if (signs)
    for (int p = tid; p < b; p += NT) signs[p] = (int64_t)sgn[p];
}

// Inverse of a triangular b×b matrix (one workgroup).  upper: invert the
// upper triangle of X; else the lower triangle (through its transpose).
__device__ __forceinline__ void tri_inverse_body(const double* __restrict__ X, int64_t ldx, double* __restrict__ Xi,
                                                 int64_t ldi, int b, int upper, int unit, double* smem) {
  constexpr int NT = 512;
  double* M = smem;
  double* d = M + b * b;
  const int tid = threadIdx.x;
  FOR_RC(rr, cc, b, NT) M[rr + cc * b] = rr < cc ? (upper ? X[rr + cc * ldx] : X[cc + rr * ldx]) : 0.0;
  for (int p = tid; p < b; p += NT) d[p] = unit ? 1.0 : 1.0 / X[p + p * ldx];
  __syncthreads();
  tri_inv_upper_lds<NT>(M, b, d, b);
  FOR_RC(rr, cc, b, NT) {
    const double v = rr <= cc ? M[rr + cc * b] : 0.0;
    if (upper) Xi[rr + cc * ldi] = v;
    else Xi[cc + rr * ldi] = v;
  }
}

__global__ void __launch_bounds__(512)
tri_inverse_kernel(const double* __restrict__ X, int64_t ldx, double* __restrict__ Xi, int64_t ldi,
                   int b, int upper, int unit) {
  extern __shared__ double smem[];
  tri_inverse_body(X, ldx, Xi, ldi, b, upper, unit, smem);
}

// Inverses of the diagonal blocks of the upper triangle of an n×n matrix,
// blk wide (the last one narrower), one workgroup per block: block p's
// inverse is stored bj×bj (ld bj) at Xi + p·blk².
__global__ void __launch_bounds__(512)
tri_inverse_diag_kernel(const double* __restrict__ X, int64_t ldx, double* __restrict__ Xi, int64_t n, int blk) {
  extern __shared__ double smem[];
  const int64_t j = (int64_t)blockIdx.x * blk;
  const int bj = (int)(n - j < blk ? n - j : blk);
  tri_inverse_body(X + j + j * ldx, ldx, Xi + (int64_t)blockIdx.x * blk * blk, bj, bj, 1, 0, smem);
}

// tau_recon (alg/shared/util.cxx:437-465), one wave per column.
__global__ void tau_recon_kernel(char s, int64_t ncol, int64_t nrow, int64_t lda,
                                 const double* __restrict__ Y, double* __restrict__ tau) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  if (i >= ncol) return;
  int64_t lo, hi;
  if (s == 'L') { lo = i + 1; hi = nrow; }
  else if (s == 'U') { lo = 0; hi = i + 1; }
  else { lo = 0; hi = nrow; }
  const double* col = Y + i * lda;
  double mx = 0.0;
  for (int64_t r = lo + lane; r < hi; r += 64) mx = fmax(mx, fabs(col[r]));
  mx = wave_max(mx);
  double nrm = 0.0;
  if (mx > 0.0) {
    double ss = 0.0;
    for (int64_t r = lo + lane; r < hi; r += 64) { const double t = col[r] / mx; ss += t * t; }
    nrm = mx * sqrt(wave_sum(ss));
  }
  if (lane == 0) {
    const double t = 2.0 / (nrm * nrm + 1.0);
    tau[i] = t == 2.0 ? 0.0 : t;
  }
}

// Packed triangles (alg/shared/util.cxx:322-406).  op: 0 pack_upper,
// 1 unpack_upper, 2 pack_lower, 3 unpack_lower.
__global__ void pack_kernel(int op, const double* __restrict__ src, double* __restrict__ dst,
                            int64_t b, int64_t lda, int has_diag) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= b * b) return;
  const int64_t rr = e % b, cc = e / b;
  if (op <= 1) {
    if (rr > cc) return;
    const int64_t p = cc * (cc + 1) / 2 + rr;
    if (op == 0) dst[p] = src[rr + cc * lda];
    else dst[rr + cc * lda] = src[p];
  } else {
    const int64_t first = cc + 1 - has_diag;
    if (rr < first) return;
    const int64_t len_before = cc * (b - 1 + has_diag) - cc * (cc - 1) / 2;
    const int64_t p = len_before + (rr - first);
    if (op == 2) dst[p] = src[rr + cc * lda];
    else dst[rr + cc * lda] = src[p];
  }
}

// Explicit unit-lower Y (ld ldy) from a panel [Y\R] (qr_2d.cxx:157-162).
__global__ void copy_unit_lower_kernel(const double* __restrict__ YR, int64_t ldyr,
                                       double* __restrict__ Y, int64_t ldy, int64_t m, int64_t b) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * b;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rr = e % m, cc = e / m;
    Y[rr + cc * ldy] = rr < cc ? 0.0 : (rr == cc ? 1.0 : YR[rr + cc * ldyr]);
  }
}

// Blocked nb×b T (LAPACK dtpqrt/dgeqrt layout) from the full b×b T.
__global__ void blocked_T_kernel(const double* __restrict__ Tf, double* __restrict__ Tb, int b, int nb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb * b) return;
  const int rr = e % nb, cc = e / nb;
  const int s = (cc / nb) * nb;
  Tb[e] = (s + rr <= cc) ? Tf[(s + rr) + cc * b] : 0.0;
}

__global__ void finite_scan_kernel(const double* __restrict__ A, int64_t lda, int64_t m, int64_t n,
                                   int* __restrict__ flag) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double v = A[(e % m) + (e / m) * lda];
    if (!isfinite(v)) *flag = 1;
  }
}

// Leaf diagonal of T: tau_l(i) = T_l(i,i) (dgeqrf tau of a leaf).
__global__ void diag_kernel(const double* __restrict__ T, int64_t ldt, double* __restrict__ tau, int b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < b) tau[p] = T[p + p * ldt];
}

// C -= X element-wise over z-batched b×b blocks is done with bgemm
// (SM_IDENTITY); no extra kernel needed.

// C_z += alpha·X_z over z-batched b×b blocks (reverse-tree C_r −= X)

// T := tril(G) with the diagonal halved (qr_2d.cxx:42-51)
__global__ void tril_halve_kernel(const double* __restrict__ G, int b, double* __restrict__ T, int64_t ldt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b * b) return;
  const int rr = e % b, cc = e / b;
  T[rr + cc * ldt] = rr < cc ? 0.0 : (rr == cc ? 0.5 * G[e] : G[e]);
}

__global__ void fill2d_kernel(double* __restrict__ D, int64_t ldd, int64_t m, int64_t n, double v) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n;
       e += (int64_t)gridDim.x * blockDim.x)
    D[(e % m) + (e / m) * ldd] = v;
}

__global__ void copy2d_kernel(const double* __restrict__ S, int64_t lds, double* __restrict__ D, int64_t ldd,
                              int64_t m, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rr = e % m, cc = e / m;
    D[rr + cc * ldd] = S[rr + cc * lds];
  }
}

// ===========================================================================
// Launchers
// ===========================================================================
template <int WAVES, int CPW, int RPL>
static int leaf_launch(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b,
                       double* R, double* T, double* tau, const BatchProb* bp = nullptr, int nprob = 0) {
  const size_t lds = sizeof(double) * ((size_t)b * b + 2 * 64 * RPL + 128);
  auto k = leaf_qr_kernel<WAVES, CPW, RPL>;
  CQR_TRY(set_max_lds(c, (const void*)k, 163840));
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(k, dim3(L), dim3(64 * WAVES), lds, c->stream, A, lda, m, L, b, R, T, tau, bp, nprob);
  prof_end(c, ev, PK_LEAF, 2.0 * m * b * b - (2.0 / 3.0) * L * (double)b * b * b, 16.0 * m * b);
  return check_launch(c, "leaf_qr");
}

// Batched: L = total leaves of the group, m = its total rows, rows = the
// largest leaf of any problem (selects the register tile).
int leaf_qr_batched(cqr_ctx* c, double* Abase, int64_t total_m, int L, int64_t rows, int b, double* R, double* T,
                    double* tau, const BatchProb* bp, int nprob) {
  if (b <= 16) return leaf_launch<4, 4, 16>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob);
  if (b <= 32) return rows <= 512 ? leaf_launch<8, 4, 8>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob)
                                  : leaf_launch<8, 4, 16>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob);
  if (b <= 64) return rows <= 256 ? leaf_launch<8, 8, 4>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob)
                                  : leaf_launch<8, 8, 8>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob);
  return rows <= 128 ? leaf_launch<16, 8, 2>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob)
                     : leaf_launch<16, 8, 4>(c, Abase, 0, total_m, L, b, R, T, tau, bp, nprob);
}

int leaf_qr(cqr_ctx* c, double* A, int64_t lda, int64_t m, int L, int b, double* R, double* T,
            double* tau) {
  // rows per lane sized to the largest leaf (the register tile is RPL·64 rows)
  const int64_t rows = (m + L - 1) / L;
  if (b <= 16) return leaf_launch<4, 4, 16>(c, A, lda, m, L, b, R, T, tau);
  if (b <= 32) return rows <= 512 ? leaf_launch<8, 4, 8>(c, A, lda, m, L, b, R, T, tau)
                                  : leaf_launch<8, 4, 16>(c, A, lda, m, L, b, R, T, tau);
  if (b <= 64) return rows <= 256 ? leaf_launch<8, 8, 4>(c, A, lda, m, L, b, R, T, tau)
                                  : leaf_launch<8, 8, 8>(c, A, lda, m, L, b, R, T, tau);
  // b > 64: sixteen waves of eight columns (109 VGPRs, four waves per SIMD
  // hide more of the column step's latency than eight waves of sixteen)
  return rows <= 128 ? leaf_launch<16, 8, 2>(c, A, lda, m, L, b, R, T, tau)
                     : leaf_launch<16, 8, 4>(c, A, lda, m, L, b, R, T, tau);
}

int64_t leaf_capacity_rows(int64_t b) {
  if (b <= 32) return 1024;
  if (b <= 64) return 512;
  return 256;
}

template <int WAVES, int CPW, int RPL>
static int node_launch(cqr_ctx* c, double* Rleaf, int n, int node0, int b, double* V2, double* T,
                       double* tau) {
  const int npairs = n / 2;
  if (npairs == 0) return CQR_OK;
  const size_t lds = sizeof(double) * ((size_t)b * b + 2 * 64 * RPL + 128);
  auto k = tree_node_kernel<WAVES, CPW, RPL>;
  CQR_TRY(set_max_lds(c, (const void*)k, 163840));
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(k, dim3(npairs), dim3(64 * WAVES), lds, c->stream, Rleaf, n, node0, b, V2, T, tau);
  prof_end(c, ev, PK_TREE, npairs * (2.0 / 3.0) * b * (double)b * b, npairs * 40.0 * b * b);
  return check_launch(c, "tree_node");
}

int tree_level(cqr_ctx* c, double* Rleaf, int n, int node0, int b, double* V2, double* T, double* tau) {
  if (b <= 16) return node_launch<4, 4, 1>(c, Rleaf, n, node0, b, V2, T, tau);
  if (b <= 32) return node_launch<8, 4, 1>(c, Rleaf, n, node0, b, V2, T, tau);
  if (b <= 64) return node_launch<8, 8, 1>(c, Rleaf, n, node0, b, V2, T, tau);
  return node_launch<8, 16, 2>(c, Rleaf, n, node0, b, V2, T, tau);
}

int bgemm(cqr_ctx* c, const BgemmArgs& g, int batch, int max_m) {
  if (batch <= 0 || g.n <= 0 || max_m <= 0) return CQR_OK;
  dim3 grid((max_m + 63) / 64, (g.n + 63) / 64, batch);
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(bgemm_kernel, grid, dim3(256), 0, c->stream, g);
  const double kk = g.k >= 0 ? (double)g.k : (double)g.pm / std::max(g.pL, 1);
  prof_end(c, ev, PK_BGEMM, 2.0 * batch * (double)max_m * g.n * kk,
           8.0 * batch * ((double)max_m * kk + kk * g.n + 2.0 * max_m * g.n));
  return check_launch(c, "bgemm");
}

static int big_lds_attr(cqr_ctx* c, const void* k) { return set_max_lds(c, k, 163840); }

int hr_core(cqr_ctx* c, const double* Q1top, int64_t ldq, const double* Rroot, double* Atop,
            int64_t lda, double* W, double* Uinv, double* Winv, double* LinvT, int64_t* signs, int b,
            double* Tt, double* M2, double* Ye, double* YT, int64_t ldy) {
  if (Tt && b != 128) return set_err(c, CQR_ERR_UNSUPPORTED, "hr_core: fused Tt needs b = 128");
  if (Ye && !(Tt && YT)) return set_err(c, CQR_ERR_INVALID, "hr_core: Ye needs Tt and YT");
  CQR_TRY(big_lds_attr(c, (const void*)hr_core_kernel));
  const size_t lds = sizeof(double) * ((size_t)b * (b + 1) + 2 * b);
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(hr_core_kernel, dim3(1), dim3(512), lds, c->stream, Q1top, ldq, Rroot, Atop,
                     lda, W, Uinv, Winv, LinvT, signs, b, Tt, M2, Ye, YT, ldy);
  const double bbb = (double)b * b * b;
  prof_end(c, ev, PK_HR, (2.0 / 3.0) * bbb + (Tt ? bbb / 3.0 : 0.0) + (M2 ? bbb : 0.0), 64.0 * b * b);
  return check_launch(c, "hr_core");
}

int hr_core_batched(cqr_ctx* c, const BatchProb* bp, int nprob, const double* Y, const double* leafR,
                    const double* nodeR, double* Abase, double* W, double* Uinv, double* Winv, double* LinvT, int b) {
  CQR_TRY(big_lds_attr(c, (const void*)hr_core_batched_kernel));
  const size_t lds = sizeof(double) * ((size_t)b * (b + 1) + 2 * b);
  hipEvent_t ev = prof_begin(c);
  hipLaunchKernelGGL(hr_core_batched_kernel, dim3(nprob), dim3(512), lds, c->stream, bp, Y, leafR, nodeR, Abase, W,
                     Uinv, Winv, LinvT, b);
  prof_end(c, ev, PK_HR, nprob * (2.0 / 3.0) * b * (double)b * b, nprob * 64.0 * b * b);
  return check_launch(c, "hr_core_batched");
}

int launch_signed_nlu(cqr_ctx* c, double* A, int64_t lda, double* R, int64_t lda_R, int64_t* signs, int b) {
  CQR_TRY(big_lds_attr(c, (const void*)signed_nlu_kernel));
  const size_t lds = sizeof(double) * ((size_t)b * (b + 1) + b);
  hipLaunchKernelGGL(signed_nlu_kernel, dim3(1), dim3(512), lds, c->stream, A, lda, R, lda_R, signs, b);
  return check_launch(c, "signed_nlu");
}

int launch_tri_inverse(cqr_ctx* c, const double* X, int64_t ldx, double* Xi, int64_t ldi, int b,
                       int upper, int unit) {
  CQR_TRY(big_lds_attr(c, (const void*)tri_inverse_kernel));
  const size_t lds = sizeof(double) * ((size_t)b * b + b);
  hipLaunchKernelGGL(tri_inverse_kernel, dim3(1), dim3(512), lds, c->stream, X, ldx, Xi, ldi, b, upper, unit);
  return check_launch(c, "tri_inverse");
}

int launch_tri_inverse_diag(cqr_ctx* c, const double* X, int64_t ldx, double* Xi, int64_t n, int blk) {
  if (n < 1 || blk < 1 || blk > 128) return set_err(c, CQR_ERR_INVALID, "tri_inverse_diag: n %lld, blk %d", (long long)n, blk);
  CQR_TRY(big_lds_attr(c, (const void*)tri_inverse_diag_kernel));
  const size_t lds = sizeof(double) * ((size_t)blk * blk + blk);
  const int64_t nb = (n + blk - 1) / blk;
  hipLaunchKernelGGL(tri_inverse_diag_kernel, dim3((unsigned)nb), dim3(512), lds, c->stream, X, ldx, Xi, n, blk);
  return check_launch(c, "tri_inverse_diag");
}



int launch_tau_recon(cqr_ctx* c, char s, int64_t ncol, int64_t nrow, int64_t lda, const double* Y,
                     double* tau) {
  if (ncol <= 0) return CQR_OK;
  hipLaunchKernelGGL(tau_recon_kernel, dim3((ncol + 3) / 4), dim3(256), 0, c->stream, s, ncol, nrow,
                     lda, Y, tau);
  return check_launch(c, "tau_recon");
}

int launch_pack(cqr_ctx* c, int op, const double* src, double* dst, int64_t b, int64_t lda, int has_diag) {
  if (b <= 0) return CQR_OK;
  hipLaunchKernelGGL(pack_kernel, dim3((b * b + 255) / 256), dim3(256), 0, c->stream, op, src, dst, b,
                     lda, has_diag);
  return check_launch(c, "pack");
}

int launch_copy_unit_lower(cqr_ctx* c, const double* YR, int64_t ldyr, double* Y, int64_t ldy,
                           int64_t m, int64_t b) {
  if (m <= 0 || b <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(4096, (m * b + 255) / 256);
  hipLaunchKernelGGL(copy_unit_lower_kernel, dim3(gx), dim3(256), 0, c->stream, YR, ldyr, Y, ldy, m, b);
  return check_launch(c, "copy_unit_lower");
}

int launch_blocked_T(cqr_ctx* c, const double* Tf, double* Tb, int b, int nb) {
  hipLaunchKernelGGL(blocked_T_kernel, dim3((nb * b + 255) / 256), dim3(256), 0, c->stream, Tf, Tb, b, nb);
  return check_launch(c, "blocked_T");
}

int launch_finite_scan(cqr_ctx* c, const double* A, int64_t lda, int64_t m, int64_t n) {
  const int gx = (int)std::min<int64_t>(4096, (m * n + 255) / 256);
  if (gx <= 0) return CQR_OK;
  hipLaunchKernelGGL(finite_scan_kernel, dim3(gx), dim3(256), 0, c->stream, A, lda, m, n, c->d_flag);
  return check_launch(c, "finite_scan");
}


int launch_tril_halve(cqr_ctx* c, const double* G, int b, double* T, int64_t ldt) {
  hipLaunchKernelGGL(tril_halve_kernel, dim3((b * b + 255) / 256), dim3(256), 0, c->stream, G, b, T, ldt);
  return check_launch(c, "tril_halve");
}

int launch_copy2d(cqr_ctx* c, const double* src, int64_t lds, double* dst, int64_t ldd, int64_t m, int64_t n) {
  if (m <= 0 || n <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(2048, (m * n + 255) / 256);
  hipLaunchKernelGGL(copy2d_kernel, dim3(gx), dim3(256), 0, c->stream, src, lds, dst, ldd, m, n);
  return check_launch(c, "copy2d");
}

int launch_fill2d(cqr_ctx* c, double* dst, int64_t ldd, int64_t m, int64_t n, double v) {
  if (m <= 0 || n <= 0) return CQR_OK;
  const int gx = (int)std::min<int64_t>(2048, (m * n + 255) / 256);
  hipLaunchKernelGGL(fill2d_kernel, dim3(gx), dim3(256), 0, c->stream, dst, ldd, m, n, v);
  return check_launch(c, "fill2d");
}

int launch_diag(cqr_ctx* c, const double* T, int64_t ldt, double* tau, int b) {
  hipLaunchKernelGGL(diag_kernel, dim3((b + 127) / 128), dim3(128), 0, c->stream, T, ldt, tau, b);
  return check_launch(c, "diag");
}

}  // namespace cqr
